// extern "C" boundary of the library (include/lsweeps_hip.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>

#include "../../include/lsweeps_hip.h"
#include "lsw_dd.h"
#include "lsw_gmres.h"
#include "lsw_input.h"
#include "lsw_model.h"
#include "lsw_output.h"
#include "lsw_serve.h"
#include "lsw_solver.h"

using namespace lsw;

struct lsw_lu {
  std::unique_ptr<LocalLinearSolverHIP> solver;
  ComplexSparseMatrix A;
};

struct lsw_dd {
  std::unique_ptr<DDCheckerboard> dd;
  GlobalIterativeSolverGMRESHIP gmres;
};

struct lsw_server {
  std::unique_ptr<Server> srv;
};

namespace {
thread_local std::string gLastError;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    gLastError.clear();
    return LSW_OK;
  } catch (const Error& e) {
    gLastError = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    gLastError = "out of host memory";
    return LSW_ERR_CAPACITY;
  } catch (const std::exception& e) {
    gLastError = e.what();
    return LSW_ERR_INTERNAL;
  } catch (...) {
    gLastError = "unknown error";
    return LSW_ERR_INTERNAL;
  }
}
}  // namespace

extern "C" {

const char* lsw_last_error(void) { return gLastError.c_str(); }
const char* lsw_version(void) { return "lsweeps-mi355x 0.1"; }

int lsw_device_count(int* count) {
  return guarded([&] {
    LSW_REQUIRE(count != nullptr, LSW_ERR_ARG, "null pointer");
    int n = 0;
    LSW_HIP(hipGetDeviceCount(&n));
    *count = n;
  });
}

int lsw_device_memory(long long* live, long long* cached, int release) {
  return guarded([&] {
    if (release) releaseDeviceCache();
    if (live) *live = totalDeviceBytes();
    if (cached) *cached = cachedDeviceBytes();
  });
}

// ------------------------------------------------------------------ LU
int lsw_lu_create(int precision, lsw_lu** out) {
  return guarded([&] {
    LSW_REQUIRE(out != nullptr, LSW_ERR_ARG, "null pointer");
    LSW_REQUIRE(precision == LSW_PREC_FP64 || precision == LSW_PREC_FP32, LSW_ERR_ARG, "unknown precision");
    auto* h = new lsw_lu();
    h->solver = std::make_unique<LocalLinearSolverHIP>(precision);
    *out = h;
  });
}

int lsw_lu_set_csc_z(lsw_lu* lu, int n, int nnz, const int* colptr, const int* rowind, const double* vals) {
  return guarded([&] {
    LSW_REQUIRE(lu && colptr && rowind && vals, LSW_ERR_ARG, "null pointer");
    LSW_REQUIRE(n > 0 && nnz >= n, LSW_ERR_ARG, "bad matrix size");
    LSW_REQUIRE(colptr[0] == 0 && colptr[n] == nnz, LSW_ERR_ARG, "bad column pointer");
    lu->A.rows = lu->A.cols = n;
    lu->A.colptr.assign(colptr, colptr + n + 1);
    lu->A.rowind.assign(rowind, rowind + nnz);
    lu->A.vals.resize(nnz);
    for (int k = 0; k < nnz; ++k) lu->A.vals[k] = Complex(vals[2 * k], vals[2 * k + 1]);
    lu->solver->setMatrix(&lu->A);
  });
}

int lsw_lu_factorize(lsw_lu* lu) {
  return guarded([&] {
    LSW_REQUIRE(lu, LSW_ERR_ARG, "null pointer");
    lu->solver->factorize();
  });
}

int lsw_lu_solve_z(lsw_lu* lu, int nrhs, const double* b, double* x) {
  return guarded([&] {
    LSW_REQUIRE(lu && b && x, LSW_ERR_ARG, "null pointer");
    lu->solver->solveMany(reinterpret_cast<const Complex*>(b), reinterpret_cast<Complex*>(x), nrhs);
  });
}

int lsw_lu_solve_dev(lsw_lu* lu, int nrhs, const void* b, void* x, void* stream) {
  return guarded([&] {
    LSW_REQUIRE(lu && b && x, LSW_ERR_ARG, "null pointer");
    lu->solver->solveDev(static_cast<const double2*>(b), static_cast<double2*>(x), nrhs,
                         static_cast<hipStream_t>(stream));
  });
}

int lsw_lu_solve_dev_noalloc(lsw_lu* lu, int nrhs, const void* b, void* x, void* stream) {
  return guarded([&] {
    LSW_REQUIRE(lu && b && x, LSW_ERR_ARG, "null pointer");
    lu->solver->solveDevNoAlloc(static_cast<const double2*>(b), static_cast<double2*>(x), nrhs,
                                static_cast<hipStream_t>(stream));
  });
}

int lsw_lu_info(lsw_lu* lu, int* nx, int* ny, long long* factor_entries, int* nodes, int* height) {
  return guarded([&] {
    LSW_REQUIRE(lu, LSW_ERR_ARG, "null pointer");
    if (nx) *nx = lu->solver->gridNx();
    if (ny) *ny = lu->solver->gridNy();
    const FactorSet* fs = lu->solver->factorSet();
    if (factor_entries) *factor_entries = fs ? fs->factorEntries() : 0;
    if (nodes) *nodes = fs ? (int)fs->tree(0).nodes.size() : 0;
    if (height) *height = fs ? fs->maxHeight() : 0;
  });
}

int lsw_lu_destroy(lsw_lu* lu) {
  return guarded([&] { delete lu; });
}

// ------------------------------------------------------------------ DD
int lsw_dd_create_config(const char* path, int id, lsw_dd** out) {
  return guarded([&] {
    LSW_REQUIRE(path && out, LSW_ERR_ARG, "null pointer");
    Input in;
    in.openFile(path);
    LSW_REQUIRE(id >= 0 && (UInt)id < in.getNumProblems(), LSW_ERR_ARG, "problem id out of range");
    Problem pr;
    pr.read(in, (UInt)id);
    auto h = std::make_unique<lsw_dd>();
    h->dd = pr.takeDD();
    h->gmres.setParent(h->dd.get());
    *out = h.release();
  });
}

static int ddCreate(const lsw_dd_params* p, lsw_dd** out, bool device) {
  return guarded([&] {
    LSW_REQUIRE(p && out, LSW_ERR_ARG, "null pointer");
    auto h = std::make_unique<lsw_dd>();
    h->dd = makeCheckerboard(*p, device);
    h->gmres.setParent(h->dd.get());
    *out = h.release();
  });
}

int lsw_dd_create(const lsw_dd_params* p, lsw_dd** out) { return ddCreate(p, out, true); }
int lsw_dd_create_host(const lsw_dd_params* p, lsw_dd** out) { return ddCreate(p, out, false); }

int lsw_dd_destroy(lsw_dd* dd) {
  return guarded([&] { delete dd; });
}

int lsw_dd_factorize(lsw_dd* dd) {
  return guarded([&] {
    LSW_REQUIRE(dd, LSW_ERR_ARG, "null pointer");
    LSW_REQUIRE(dd->dd->deviceReady(), LSW_ERR_STATE, "DD was created without device state");
    dd->dd->factorize();
  });
}

int lsw_dd_sizes(lsw_dd* dd, int* Gx, int* Gy, int* ncells, double* mesh_size) {
  return guarded([&] {
    LSW_REQUIRE(dd, LSW_ERR_ARG, "null pointer");
    if (Gx) *Gx = dd->dd->Gx();
    if (Gy) *Gy = dd->dd->Gy();
    if (ncells) *ncells = (int)dd->dd->getNCells();
    if (mesh_size) *mesh_size = dd->dd->getMeshSize();
  });
}

int lsw_dd_global_coords(lsw_dd* dd, double* xs, double* ys) {
  return guarded([&] {
    LSW_REQUIRE(dd && xs && ys, LSW_ERR_ARG, "null pointer");
    for (int i = 0; i < dd->dd->Gx(); ++i) xs[i] = dd->dd->gx(i);
    for (int j = 0; j < dd->dd->Gy(); ++j) ys[j] = dd->dd->gy(j);
  });
}

int lsw_dd_cell_info(lsw_dd* dd, int cell, int* info) {
  return guarded([&] {
    LSW_REQUIRE(dd && info, LSW_ERR_ARG, "null pointer");
    LSW_REQUIRE(cell >= 0 && cell < (int)dd->dd->getNCells(), LSW_ERR_ARG, "cell index out of range");
    const CellGeom& c = dd->dd->cellGeom(cell);
    info[0] = c.I;
    info[1] = c.J;
    info[2] = c.nx;
    info[3] = c.ny;
    info[4] = c.ox;
    info[5] = c.oy;
    for (int f = 0; f < 2; ++f)
      for (int k = 0; k < 8; ++k) {
        info[6 + f * 8 + k] = c.vlr[f][k];
        info[22 + f * 8 + k] = c.vud[f][k];
      }
  });
}

int lsw_dd_assemble_rhs(lsw_dd* dd, int nsrc, const double* src_xy, double* b) {
  return guarded([&] {
    LSW_REQUIRE(dd && b && (nsrc == 0 || src_xy), LSW_ERR_ARG, "null pointer");
    std::vector<DoubleVector> s(nsrc);
    for (int k = 0; k < nsrc; ++k) s[k] = {src_xy[2 * k], src_xy[2 * k + 1]};
    const long long n = dd->dd->N();
    double2* d = nullptr;
    LSW_HIP(devMalloc(&d, n * sizeof(double2)));
    try {
      dd->dd->assembleRHSDev(s, d, dd->dd->stream());
      LSW_HIP(hipMemcpy(b, d, n * sizeof(double2), hipMemcpyDeviceToHost));
    } catch (...) {
      (void)devFree(d);
      throw;
    }
    (void)devFree(d);
  });
}

int lsw_dd_assemble_rhs_dev(lsw_dd* dd, int nsrc, const double* src_xy, void* b, void* stream) {
  return guarded([&] {
    LSW_REQUIRE(dd && b && (nsrc == 0 || src_xy), LSW_ERR_ARG, "null pointer");
    std::vector<DoubleVector> s(nsrc);
    for (int k = 0; k < nsrc; ++k) s[k] = {src_xy[2 * k], src_xy[2 * k + 1]};
    dd->dd->assembleRHSDev(s, (double2*)b, (hipStream_t)stream);
  });
}

namespace {
void hostApply(lsw_dd* dd, const double* x, double* y, bool precond) {
  const long long n = dd->dd->N();
  double2 *dx = nullptr, *dy = nullptr;
  LSW_HIP(devMalloc(&dx, n * sizeof(double2)));
  LSW_HIP(devMalloc(&dy, n * sizeof(double2)));
  try {
    LSW_HIP(uploadSync(dx, x, n * sizeof(double2)));
    if (precond)
      dd->dd->applyPrecondDev(dx, dy, dd->dd->stream(), 1, false);
    else
      dd->dd->applyADev(dx, dy, dd->dd->stream());
    LSW_HIP(hipStreamSynchronize(dd->dd->stream()));
    LSW_HIP(hipMemcpy(y, dy, n * sizeof(double2), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)devFree(dx);
    (void)devFree(dy);
    throw;
  }
  (void)devFree(dx);
  (void)devFree(dy);
}
}  // namespace

int lsw_dd_apply_precond(lsw_dd* dd, const double* x, double* y) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y, LSW_ERR_ARG, "null pointer");
    hostApply(dd, x, y, true);
  });
}

int lsw_dd_apply_precond_dev(lsw_dd* dd, const void* x, void* y, void* stream) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y, LSW_ERR_ARG, "null pointer");
    dd->dd->applyPrecondDev((const double2*)x, (double2*)y, (hipStream_t)stream);
  });
}

int lsw_dd_apply_precond_block(lsw_dd* dd, int nrhs, const double* x, double* y) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y && nrhs > 0, LSW_ERR_ARG, "null pointer or bad nrhs");
    const long long n = dd->dd->N() * nrhs;
    double2 *dx = nullptr, *dy = nullptr;
    LSW_HIP(devMalloc(&dx, n * sizeof(double2)));
    LSW_HIP(devMalloc(&dy, n * sizeof(double2)));
    try {
      LSW_HIP(uploadSync(dx, x, n * sizeof(double2)));
      dd->dd->applyPrecondDev(dx, dy, dd->dd->stream(), nrhs, false);
      LSW_HIP(hipStreamSynchronize(dd->dd->stream()));
      LSW_HIP(hipMemcpy(y, dy, n * sizeof(double2), hipMemcpyDeviceToHost));
    } catch (...) {
      (void)devFree(dx);
      (void)devFree(dy);
      throw;
    }
    (void)devFree(dx);
    (void)devFree(dy);
  });
}

int lsw_dd_apply_precond_block_dev(lsw_dd* dd, int nrhs, const void* x, void* y, void* stream) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y && nrhs > 0, LSW_ERR_ARG, "null pointer or bad nrhs");
    dd->dd->applyPrecondDev((const double2*)x, (double2*)y, (hipStream_t)stream, nrhs);
  });
}

int lsw_dd_apply_a(lsw_dd* dd, const double* x, double* y) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y, LSW_ERR_ARG, "null pointer");
    hostApply(dd, x, y, false);
  });
}

int lsw_dd_apply_a_dev(lsw_dd* dd, const void* x, void* y, void* stream) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y, LSW_ERR_ARG, "null pointer");
    dd->dd->applyADev((const double2*)x, (double2*)y, (hipStream_t)stream);
  });
}

int lsw_dd_stats(lsw_dd* dd, long long* solves, double* bytes, long long* factor_bytes) {
  return guarded([&] {
    LSW_REQUIRE(dd, LSW_ERR_ARG, "null pointer");
    if (solves) *solves = dd->dd->localSolvesPerPrecond();
    if (bytes) *bytes = dd->dd->solveBytesPerPrecond();
    if (factor_bytes) *factor_bytes = dd->dd->factors().factorized() ? dd->dd->factors().factorBytes() : 0;
  });
}

int lsw_dd_profile_precond(lsw_dd* dd, const void* x, void* y, double* out) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y && out, LSW_ERR_ARG, "null pointer");
    SolveProfile prof;
    dd->dd->applyPrecondProfiled((const double2*)x, (double2*)y, prof);
    out[0] = prof.fwdMs;
    out[1] = prof.fwdLaunches;
    out[2] = prof.fwdBytes;
    out[3] = prof.bwdMs;
    out[4] = prof.bwdLaunches;
    out[5] = prof.bwdBytes;
    out[6] = prof.fwdFlops;
    out[7] = prof.bwdFlops;
    out[8] = prof.idealMs;
  });
}

int lsw_dd_profile_precond_block(lsw_dd* dd, int nrhs, const void* x, void* y, double* out) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && y && out, LSW_ERR_ARG, "null pointer");
    SolveProfile prof;
    LSW_REQUIRE(nrhs >= 1, LSW_ERR_ARG, "nrhs must be positive");
    dd->dd->applyPrecondProfiled((const double2*)x, (double2*)y, prof, nrhs);
    out[0] = prof.fwdMs;
    out[1] = prof.fwdLaunches;
    out[2] = prof.fwdBytes;
    out[3] = prof.bwdMs;
    out[4] = prof.bwdLaunches;
    out[5] = prof.bwdBytes;
    out[6] = prof.fwdFlops;
    out[7] = prof.bwdFlops;
    out[8] = prof.idealMs;
  });
}

int lsw_dd_set_graphs(lsw_dd* dd, int on) {
  return guarded([&] {
    LSW_REQUIRE(dd, LSW_ERR_ARG, "null pointer");
    dd->dd->setGraphs(on != 0);
  });
}

int lsw_dd_autotune(lsw_dd* dd, int nrhs, double* out) {
  return guarded([&] {
    LSW_REQUIRE(dd && out, LSW_ERR_ARG, "null pointer");
    const auto r = dd->dd->autotune(nrhs);
    out[0] = r.defaultMs;
    out[1] = r.bestMs;
    out[2] = (double)r.tried;
    out[3] = (double)r.shape.latencyWaves;
    out[4] = (double)r.shape.wideBudget;
    out[5] = (double)r.shape.wideCap;
    out[6] = r.shape.mfma ? 1.0 : 0.0;
    out[7] = (double)r.shape.latencyRows;
    out[8] = (double)r.shape.mfmaWideCols;
    out[9] = r.shape.mfmaPair ? 1.0 : 0.0;
    out[10] = r.shape.mfmaGroup ? 1.0 : 0.0;
  });
}

int lsw_dd_sync(lsw_dd* dd) {
  return guarded([&] {
    LSW_REQUIRE(dd, LSW_ERR_ARG, "null pointer");
    LSW_HIP(hipStreamSynchronize(dd->dd->stream()));
  });
}

// ------------------------------------------------------------------ GMRES
namespace {
void configure(lsw_dd* dd, const lsw_gmres_params* p) {
  LSW_REQUIRE(p->mode == 0 || p->mode == 1, LSW_ERR_ARG, "unknown GMRES mode");
  LSW_REQUIRE(p->max_iter >= 0 && p->max_iter <= 1000, LSW_ERR_ARG, "bad MaxIter");
  LSW_REQUIRE(p->tol >= 0, LSW_ERR_ARG, "bad Tol");
  dd->gmres.setMode(p->mode);
  dd->gmres.setTol(p->tol);
  dd->gmres.setMaxIter(p->max_iter);
  dd->gmres.setRestart(p->restart > 0 ? p->restart : 30);
  dd->gmres.setVerbose(p->verbose != 0);
}
void fill(const GmresInfo& in, lsw_gmres_result* r) {
  if (!r) return;
  r->iterations = in.iterations;
  r->precond_applies = in.preconditionerApplies;
  r->matvecs = in.matvecs;
  r->estimate = in.estimate;
  r->true_residual = in.trueResidual;
}
}  // namespace

int lsw_gmres_solve(lsw_dd* dd, const lsw_gmres_params* p, const double* b, double* x, lsw_gmres_result* res,
                    double* log, int log_cap) {
  return guarded([&] {
    LSW_REQUIRE(dd && p && b && x, LSW_ERR_ARG, "null pointer");
    configure(dd, p);
    const long long n = dd->dd->N();
    double2 *db = nullptr, *dx = nullptr;
    LSW_HIP(devMalloc(&db, n * sizeof(double2)));
    LSW_HIP(devMalloc(&dx, n * sizeof(double2)));
    GmresInfo info;
    try {
      LSW_HIP(uploadSync(db, b, n * sizeof(double2)));
      dd->gmres.solveDev(db, dx, &info, dd->dd->stream());
      LSW_HIP(hipStreamSynchronize(dd->dd->stream()));
      LSW_HIP(hipMemcpy(x, dx, n * sizeof(double2), hipMemcpyDeviceToHost));
    } catch (...) {
      (void)devFree(db);
      (void)devFree(dx);
      throw;
    }
    (void)devFree(db);
    (void)devFree(dx);
    fill(info, res);
    if (log)
      for (int k = 0; k < log_cap && k < (int)info.log.size(); ++k) log[k] = info.log[k];
  });
}

int lsw_gmres_solve_dev(lsw_dd* dd, const lsw_gmres_params* p, const void* b, void* x, void* stream,
                        lsw_gmres_result* res) {
  return guarded([&] {
    LSW_REQUIRE(dd && p && b && x, LSW_ERR_ARG, "null pointer");
    configure(dd, p);
    GmresInfo info;
    dd->gmres.solveDev((const double2*)b, (double2*)x, &info, (hipStream_t)stream);
    fill(info, res);
  });
}

int lsw_gmres_solve_block_dev(lsw_dd* dd, const lsw_gmres_params* p, int nrhs, const void* B, void* X, void* stream,
                              lsw_gmres_result* res) {
  return guarded([&] {
    LSW_REQUIRE(dd && p && B && X && nrhs > 0, LSW_ERR_ARG, "null pointer or bad nrhs");
    configure(dd, p);
    std::vector<GmresInfo> infos;
    dd->gmres.solveBlockDev((const double2*)B, (double2*)X, nrhs, &infos, (hipStream_t)stream);
    if (res)
      for (int r = 0; r < nrhs; ++r) fill(infos[r], res + r);
  });
}

int lsw_gmres_solve_block(lsw_dd* dd, const lsw_gmres_params* p, int nrhs, const double* B, double* X,
                          lsw_gmres_result* res) {
  return guarded([&] {
    LSW_REQUIRE(dd && p && B && X && nrhs > 0, LSW_ERR_ARG, "null pointer or bad nrhs");
    configure(dd, p);
    const long long n = dd->dd->N() * nrhs;
    double2 *db = nullptr, *dx = nullptr;
    LSW_HIP(devMalloc(&db, n * sizeof(double2)));
    LSW_HIP(devMalloc(&dx, n * sizeof(double2)));
    std::vector<GmresInfo> infos;
    try {
      LSW_HIP(uploadSync(db, B, n * sizeof(double2)));
      dd->gmres.solveBlockDev(db, dx, nrhs, &infos, dd->dd->stream());
      LSW_HIP(hipStreamSynchronize(dd->dd->stream()));
      LSW_HIP(hipMemcpy(X, dx, n * sizeof(double2), hipMemcpyDeviceToHost));
    } catch (...) {
      (void)devFree(db);
      (void)devFree(dx);
      throw;
    }
    (void)devFree(db);
    (void)devFree(dx);
    if (res)
      for (int r = 0; r < nrhs; ++r) fill(infos[r], res + r);
  });
}

int lsw_dd_assemble_rhs_block_dev(lsw_dd* dd, int nsrc, const double* src_xy, void* b, void* stream) {
  return guarded([&] {
    LSW_REQUIRE(dd && b && src_xy && nsrc > 0, LSW_ERR_ARG, "null pointer or bad nsrc");
    const long long n = dd->dd->N();
    for (int k = 0; k < nsrc; ++k)
      dd->dd->assembleRHSDev({{src_xy[2 * k], src_xy[2 * k + 1]}}, (double2*)b + k * n, (hipStream_t)stream);
  });
}

// ------------------------------------------------------------ Input/Problem
int lsw_config_num_problems(const char* path, int* n) {
  return guarded([&] {
    LSW_REQUIRE(path && n, LSW_ERR_ARG, "null pointer");
    Input in;
    in.openFile(path);
    *n = (int)in.getNumProblems();
  });
}

int lsw_config_problem(const char* path, int id, lsw_problem_spec* spec, double* src_xy, int src_cap) {
  return guarded([&] {
    LSW_REQUIRE(path && spec, LSW_ERR_ARG, "null pointer");
    LSW_REQUIRE(id >= 0, LSW_ERR_ARG, "Invalid problem ID");
    Input in;
    in.openFile(path);
    const ProblemSpec sp = in.readProblem((UInt)id);
    std::memset(spec, 0, sizeof(*spec));
    lsw_dd_params& p = spec->dd;
    p.domain[0] = sp.domain[0][0];
    p.domain[1] = sp.domain[0][1];
    p.domain[2] = sp.domain[1][0];
    p.domain[3] = sp.domain[1][1];
    p.ncells_x = (int)sp.nCellsX;
    p.ncells_y = (int)sp.nCellsY;
    p.nx = (int)sp.NX;
    p.ny = (int)sp.NY;
    p.omega = sp.omega;
    p.min_wavelengths_in_pml = sp.minWavelengthsInPML;
    p.order = (int)sp.order;
    p.precision = sp.precision;
    p.medium_kind = sp.mediumType == "CSV" ? 1 : 0;
    if (sp.mediumType == "CSV") {
      p.vel_domain[0] = sp.csvDomain[0][0];
      p.vel_domain[1] = sp.csvDomain[0][1];
      p.vel_domain[2] = sp.csvDomain[1][0];
      p.vel_domain[3] = sp.csvDomain[1][1];
    }
    spec->csv_medium = p.medium_kind;
    spec->solver_mode = sp.solverType == "FGMRESHIP" ? 1 : 0;
    spec->tol = sp.tol;
    spec->max_iter = (int)sp.maxIter;
    spec->restart = (int)sp.restart;
    spec->verbose = sp.verbose ? 1 : 0;
    spec->nsources = (int)sp.sources.size();
    if (src_xy)
      for (int k = 0; k < src_cap && k < (int)sp.sources.size(); ++k) {
        src_xy[2 * k] = sp.sources[k][0];
        src_xy[2 * k + 1] = sp.sources[k][1];
      }
  });
}

int lsw_problem_run(const char* path, int id, int verbose, double* x, long long x_cap, lsw_problem_result* res) {
  return guarded([&] {
    LSW_REQUIRE(path, LSW_ERR_ARG, "null pointer");
    Input in;
    in.openFile(path);
    Problem pb;
    pb.read(in, (UInt)id);
    const ProblemResult r = pb.solve(verbose != 0);
    if (res) {
      res->setup_seconds = r.setupSeconds;
      res->factorize_seconds = r.factorizeSeconds;
      res->gmres_seconds = r.gmresSeconds;
      res->iterations = r.iterations;
      res->estimate = r.estimate;
      res->true_residual = r.trueResidual;
      res->dofs = r.dofs;
      res->Gx = pb.dd().Gx();
      res->Gy = pb.dd().Gy();
    }
    if (x) {
      LSW_REQUIRE(x_cap >= (long long)pb.solution().size(), LSW_ERR_CAPACITY, "solution buffer too small");
      std::memcpy(x, pb.solution().data(), pb.solution().size() * sizeof(Complex));
    }
  });
}

// ------------------------------------------------------------------ output
int lsw_dd_write_vtk(lsw_dd* dd, const double* x, const char* path) {
  return guarded([&] {
    LSW_REQUIRE(dd && x && path, LSW_ERR_ARG, "null pointer");
    const size_t n = (size_t)dd->dd->Gx() * dd->dd->Gy();
    std::vector<Complex> v(reinterpret_cast<const Complex*>(x), reinterpret_cast<const Complex*>(x) + n);
    writeVTK(std::string(path) + ".vtk", *dd->dd, v);
    writeGrid(std::string(path) + ".lswgrid", *dd->dd, v);
  });
}

// ------------------------------------------------------------------ serving
int lsw_server_create(const lsw_server_options* o, lsw_server** out) {
  return guarded([&] {
    LSW_REQUIRE(o && out, LSW_ERR_ARG, "null pointer");
    ServeOptions so;
    so.budgetBytes = o->budget_bytes;
    so.slots = o->slots;
    so.restart = o->restart;
    so.tol = o->tol;
    so.maxIter = o->max_iter;
    so.precision = o->precision;
    so.workers = o->workers;
    LSW_REQUIRE(o->order == LSW_ORDER_OLDEST || o->order == LSW_ORDER_BATCH, LSW_ERR_ARG, "unknown server order");
    so.batchOrder = o->order == LSW_ORDER_BATCH;
    auto h = std::make_unique<lsw_server>();
    h->srv = std::make_unique<Server>(so);
    *out = h.release();
  });
}

int lsw_server_destroy(lsw_server* s) {
  return guarded([&] { delete s; });
}

int lsw_server_add_problem(lsw_server* s, const lsw_dd_params* p, int* problem, int* Gx, int* Gy) {
  return guarded([&] {
    LSW_REQUIRE(s && p && problem, LSW_ERR_ARG, "null pointer");
    *problem = s->srv->addProblem(*p, Gx, Gy);
  });
}

int lsw_server_problem_bytes(lsw_server* s, int problem, double* steady, double* peak) {
  return guarded([&] {
    LSW_REQUIRE(s, LSW_ERR_ARG, "null pointer");
    s->srv->problemBytes(problem, steady, peak);
  });
}

int lsw_server_submit(lsw_server* s, int problem, int nsrc, const double* src_xy, double* x, long long* request) {
  return guarded([&] {
    LSW_REQUIRE(s && src_xy && request, LSW_ERR_ARG, "null pointer");
    *request = s->srv->submit(problem, nsrc, src_xy, x);
  });
}

int lsw_server_wait(lsw_server* s, long long request, lsw_request_result* res) {
  return guarded([&] {
    LSW_REQUIRE(s && res, LSW_ERR_ARG, "null pointer");
    const ServeResult r = s->srv->wait(request);
    res->status = r.status;
    res->problem = r.problem;
    res->nsrc = r.nsrc;
    res->max_iterations = r.maxIterations;
    res->max_true_residual = r.maxTrueResidual;
    res->submit_s = r.submitS;
    res->start_s = r.startS;
    res->end_s = r.endS;
    std::memset(res->message, 0, sizeof(res->message));
    std::strncpy(res->message, r.message.c_str(), sizeof(res->message) - 1);
  });
}

int lsw_server_get_stats(lsw_server* s, lsw_server_stats* st) {
  return guarded([&] {
    LSW_REQUIRE(s && st, LSW_ERR_ARG, "null pointer");
    const ServeStats v = s->srv->stats();
    st->requests = v.requests;
    st->completed = v.completed;
    st->solves = v.solves;
    st->loads = v.loads;
    st->evictions = v.evictions;
    st->steps = v.steps;
    st->reserved_bytes = v.reservedBytes;
    st->peak_reserved_bytes = v.peakReservedBytes;
    st->resident = v.resident;
    st->merged_steps = v.mergedSteps;
    st->max_active = v.maxActive;
    st->setup_s = v.setupS;
    st->factor_s = v.factorS;
    st->block_s = v.blockS;
    st->step_s = v.stepS;
  });
}

}  // extern "C"
