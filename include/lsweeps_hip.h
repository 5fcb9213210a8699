/*
 * C-ABI of the MI355X L-sweeps preconditioned Helmholtz solve path.
 *
 * Plain pointers and sizes only.  Complex vectors are interleaved
 * (re, im) doubles.  Host-pointer entry points (suffix _z or none) are
 * synchronous; _dev entry points take device pointers and a hipStream_t
 * passed as void* and are asynchronous.  NULL means the object's own
 * non-blocking stream, which is ordered neither with the legacy default
 * stream nor with any other stream: a caller that fills inputs on another
 * stream synchronises (or passes that stream) before the call.
 * Every function returns 0 on success or a nonzero status (LSW_ERR_*); the
 * message of the last failure on the calling thread is lsw_last_error().
 * All pointers are borrowed for the duration of the call.  Calls on one
 * object must not overlap.
 *
 * Global grid vectors: Gx*Gy entries, index gi*Gy + gj (x-major, y fastest),
 * DOF (gi, gj) at (x0 + (gi+1) hx, y0 + (gj+1) hy) of the PML-extended
 * domain (the numbering of AssemblerHelmholtz2DFD::getDOFPts,
 * src/AssemblerHelmholtz2DFD.cpp:147-181, for a 1x1 decomposition).
 */
#ifndef LSWEEPS_HIP_H
#define LSWEEPS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define LSW_OK 0
#define LSW_ERR_ARG 1
#define LSW_ERR_HIP 2
#define LSW_ERR_STATE 3
#define LSW_ERR_NUMERIC 4
#define LSW_ERR_CAPACITY 5
#define LSW_ERR_UNSUPPORTED 6
#define LSW_ERR_INTERNAL 7

#define LSW_PREC_FP64 0
#define LSW_PREC_FP32 1

typedef struct lsw_lu lsw_lu;
typedef struct lsw_dd lsw_dd;

/* Message of the last failing call on this thread ("" if none). */
const char* lsw_last_error(void);
/* Library version string. */
const char* lsw_version(void);
/* Number of visible HIP devices; selects nothing. */
int lsw_device_count(int* count);
/* Device memory the library holds: live = in use by its objects, cached =
 * freed blocks it keeps for reuse (either pointer may be NULL).  No
 * counterpart in the reference (the CPU code has no device memory): the
 * library keeps freed device blocks for the life of the process, and a
 * caller that shares the device with another allocator (torch's, say) hands
 * them back with release != 0 (after a device synchronisation). */
int lsw_device_memory(long long* live, long long* cached, int release);

/* ---- LocalLinearSolver (include/LocalLinearSolver.h:27-40) -------------
 * Replaces LocalLinearSolverEigenSparseLU (src/LocalLinearSolverEigenSparseLU.cpp:33-76):
 * setMatrix (:33-38), factorize (:40-62), solve (:65-76). */
int lsw_lu_create(int precision, lsw_lu** out);
/* CSC, int32 indices, n x n with nnz entries (Eigen column-major layout,
 * include/typedef.h:64-73).  The matrix must be the 5-point operator of an
 * nx x ny grid (index i*ny+j); the arrays are copied. */
int lsw_lu_set_csc_z(lsw_lu* lu, int n, int nnz, const int* colptr, const int* rowind, const double* vals);
int lsw_lu_factorize(lsw_lu* lu);
/* b, x: n x nrhs column-major, interleaved complex (host). */
int lsw_lu_solve_z(lsw_lu* lu, int nrhs, const double* b, double* x);
/* The same on device memory: b, x are n x nrhs column-major interleaved
 * complex doubles in HBM (b == x allowed), queued on `stream` (NULL: the
 * solver's own non-blocking stream, not ordered with other streams; wait
 * with a stream synchronisation).  The solve of
 * LocalLinearSolverEigenSparseLU::solve (src/LocalLinearSolverEigenSparseLU.cpp:65-76)
 * without the host round trip. */
int lsw_lu_solve_dev(lsw_lu* lu, int nrhs, const void* b, void* x, void* stream);
/* lsw_lu_solve_dev that allocates no device memory (no hipMalloc/hipFree,
 * no allocating library call): for callers that cannot allocate, such as
 * code under HIP graph capture or tight serving loops.  No counterpart in
 * the reference.  The solve's scratch (work lists and solve buffers) lives
 * in the library's fixed device array of 64 MiB per device, held until the
 * process ends; a solve that needs more fails with LSW_ERR_CAPACITY.  Every
 * call on one device shares that device's array: calls that overlap on one
 * device race.  The solver must be factorised; the results equal
 * lsw_lu_solve_dev's. */
int lsw_lu_solve_dev_noalloc(lsw_lu* lu, int nrhs, const void* b, void* x, void* stream);
/* nx, ny of the recovered grid, complex factor entries, tree nodes, height. */
int lsw_lu_info(lsw_lu* lu, int* nx, int* ny, long long* factor_entries, int* nodes, int* height);
int lsw_lu_destroy(lsw_lu* lu);

/* ---- DD / DDCheckerboard (include/DD.h:40-113, src/DDCheckerboard.cpp:87-645) ----
 * Parameters of Input::readCheckerboardDD / readAssemblerHelmholtz2DFD
 * (src/Input.cpp:470-623, 795-857). */
typedef struct lsw_dd_params {
  double domain[4];              /* ax, bx, ay, by */
  int ncells_x, ncells_y;        /* NCellsX (rows, y), NCellsY (columns, x) */
  int nx, ny;                    /* NX, NY: elements inside one cell */
  double omega;                  /* Omega */
  double min_wavelengths_in_pml; /* MinWavelengthsInPML */
  int order;                     /* Order (2) */
  int precision;                 /* LSW_PREC_FP64 or LSW_PREC_FP32 local factors/sweeps */
  int medium_kind;               /* 0: "Constant" (m = 1), 1: "CSV" velocities */
  const float* velocity;         /* medium_kind 1: vel_rows x vel_cols, row 0 = top */
  int vel_rows, vel_cols;
  double vel_domain[4];          /* file domain ax, bx, ay, by */
  int geometry;                  /* 0: the reference's spacing rule (needs NCellsX*(NX+1) ==
                                    NCellsY*(NY+1)); 1: per-axis spacing (rectangular domains) */
} lsw_dd_params;

int lsw_dd_create(const lsw_dd_params* p, lsw_dd** out);
/* The DD of problem `id` of a reference-format input file, built from what
 * Input::readCheckerboardDD reads (src/Input.cpp:470-623, through
 * Input::readProblemDD and Problem::read, src/Problem.cpp:70-78): domain,
 * board, Omega, PML, the "Constant" or "CSV" squared slowness (the CSV file
 * is read) and CheckerboardDD.Precision.  Not yet factorised. */
int lsw_dd_create_config(const char* path, int id, lsw_dd** out);
/* Geometry only (no device memory): sizes, coordinates and cell_info work;
 * every compute entry point returns LSW_ERR_STATE. */
int lsw_dd_create_host(const lsw_dd_params* p, lsw_dd** out);
int lsw_dd_destroy(lsw_dd* dd);
/* DDCheckerboard::factorize (src/DDCheckerboard.cpp:648-665). */
int lsw_dd_factorize(lsw_dd* dd);
/* Global grid size, number of cells, DD mesh size (getMeshSize, :3536-3561). */
int lsw_dd_sizes(lsw_dd* dd, int* Gx, int* Gy, int* ncells, double* mesh_size);
/* Global DOF coordinates (xs: Gx entries, ys: Gy entries). */
int lsw_dd_global_coords(lsw_dd* dd, double* xs, double* ys);
/* Per-cell geometry: info[0..5] = I, J, nx, ny, ox, oy; info[6..21] = LR
 * 7-set starts [unshifted 8, shifted 8]; info[22..37] = UD starts. */
int lsw_dd_cell_info(lsw_dd* dd, int cell, int* info);
/* Sum of point sources (RHSSrcPts, src/RHSSrcPts.cpp:87-112), global grid. */
int lsw_dd_assemble_rhs(lsw_dd* dd, int nsrc, const double* src_xy, double* b);
int lsw_dd_assemble_rhs_dev(lsw_dd* dd, int nsrc, const double* src_xy, void* b, void* stream);
/* y = M^{-1} x: DDCheckerboard::applyPrecond (src/DDCheckerboard.cpp:687-739). */
int lsw_dd_apply_precond(lsw_dd* dd, const double* x, double* y);
int lsw_dd_apply_precond_dev(lsw_dd* dd, const void* x, void* y, void* stream);
/* Block form: x, y hold nrhs consecutive global vectors.  The local solves of
 * each front share one read of the cell's factors among the right-hand sides. */
int lsw_dd_apply_precond_block(lsw_dd* dd, int nrhs, const double* x, double* y);
int lsw_dd_apply_precond_block_dev(lsw_dd* dd, int nrhs, const void* x, void* y, void* stream);
/* y = A x: DDCheckerboard::applySystemMatrix (src/DDCheckerboard.cpp:2611-2635). */
int lsw_dd_apply_a(lsw_dd* dd, const double* x, double* y);
int lsw_dd_apply_a_dev(lsw_dd* dd, const void* x, void* y, void* stream);
/* Local solves per applyPrecond, algorithmic bytes of those solves, factor bytes. */
int lsw_dd_stats(lsw_dd* dd, long long* solves_per_precond, double* solve_bytes_per_precond,
                 long long* factor_bytes);
/* on != 0: the _dev preconditioner entry points capture their launch
 * sequence per (x, y, nrhs, stream) into a HIP graph once and replay it
 * (off by default).  While a graph is being captured no other thread may
 * issue legacy-stream work.  No counterpart in the reference. */
int lsw_dd_set_graphs(lsw_dd* dd, int on);
/* Launch-shape autotuning for blocks of nrhs right-hand sides: times M^-1
 * under the default shape and under alternatives (column-split thresholds
 * and caps, matrix cores on/off), one choice at a time, and keeps the
 * fastest for that block width (synchronous).  out[0..10] = default ms, best
 * ms per application, shapes tried, chosen latency-wave threshold, wide
 * budget, wide cap, matrix cores (0/1), latency row limit, fewest columns
 * of a matrix-core 2-wavefront split (2^30: off), paired matrix-core
 * entries (0/1), grouped matrix-core entries (0/1).  No counterpart in the
 * reference. */
int lsw_dd_autotune(lsw_dd* dd, int nrhs, double* out);
/* Wait for all work queued on the DD's own stream. */
int lsw_dd_sync(lsw_dd* dd);
/* One applyPrecond on device vectors with every batched local-solve launch
 * timed by HIP events on the DD's stream (synchronous; not for timed
 * regions).  out[0..8] = forward: summed ms, launches, algorithmic bytes;
 * backward: summed ms, launches, algorithmic bytes; forward and backward
 * algorithmic flops; the summed roofline time in ms (per launch the larger of
 * its bytes at 8 TB/s and its flops at the matrix-core peak of the solve
 * precision).  No counterpart in the reference: it serves the roofline
 * measurement. */
int lsw_dd_profile_precond(lsw_dd* dd, const void* x, void* y, double* out);
/* The same for a block of nrhs right-hand sides (x, y: nrhs consecutive
 * global vectors on the device).  No counterpart in the reference. */
int lsw_dd_profile_precond_block(lsw_dd* dd, int nrhs, const void* x, void* y, double* out);

/* ---- GlobalIterativeSolver (include/GlobalIterativeSolver.h:36-48) -------
 * mode 0: the reference GMRES (src/GlobalIterativeSolverGMRES.cpp:65-348),
 * tol on its logged estimate, max_iter a hard cap;
 * mode 1: flexible right-preconditioned GMRES(restart), tol on the true
 * relative residual. */
typedef struct lsw_gmres_params {
  int mode;
  double tol;
  int max_iter;
  int restart;
  int verbose;
} lsw_gmres_params;

typedef struct lsw_gmres_result {
  int iterations;
  int precond_applies;
  int matvecs;
  double estimate;      /* last logged residual estimate */
  double true_residual; /* ||b - A x|| / ||b|| on the global operator, FP64 */
} lsw_gmres_result;

/* b, x: global grid vectors on the host; log (optional) receives up to
 * log_cap logged estimates. */
int lsw_gmres_solve(lsw_dd* dd, const lsw_gmres_params* p, const double* b, double* x,
                    lsw_gmres_result* res, double* log, int log_cap);
int lsw_gmres_solve_dev(lsw_dd* dd, const lsw_gmres_params* p, const void* b, void* x, void* stream,
                        lsw_gmres_result* res);
/* Flexible GMRES for nrhs right-hand sides sharing the factors (B, X: nrhs
 * consecutive global vectors; res: nrhs results).  Every Arnoldi step applies
 * the preconditioner to the block of right-hand sides whose estimate has not
 * converged yet (a converged one leaves the block: its precond_applies equal
 * its Arnoldi steps).  No
 * counterpart in the reference, which sums all sources into one right-hand
 * side (src/RHSSrcPts.cpp:93-94). */
int lsw_gmres_solve_block(lsw_dd* dd, const lsw_gmres_params* p, int nrhs, const double* B, double* X,
                          lsw_gmres_result* res);
int lsw_gmres_solve_block_dev(lsw_dd* dd, const lsw_gmres_params* p, int nrhs, const void* B, void* X,
                              void* stream, lsw_gmres_result* res);
/* Point-source right-hand sides, one per source: b holds nsrc global vectors. */
int lsw_dd_assemble_rhs_block_dev(lsw_dd* dd, int nsrc, const double* src_xy, void* b, void* stream);

/* ---- Input / Problem (include/Input.h:44-76, include/Problem.h:44-47) -----
 * The reference's libconfig input file (src/Input.cpp) and driver
 * (src/Problem.cpp:94-158). */
typedef struct lsw_problem_spec {
  lsw_dd_params dd;        /* velocity is NULL for CSV media (file named in the config) */
  int csv_medium;          /* 1 when SquaredSlowness Type = "CSV" */
  int solver_mode;         /* 0 reference GMRES ("GMRES"/"GMRESHIP"), 1 flexible ("FGMRESHIP") */
  double tol;
  int max_iter;
  int restart;
  int verbose;
  int nsources;
} lsw_problem_spec;

typedef struct lsw_problem_result {
  double setup_seconds;
  double factorize_seconds; /* "Time to factorize system" (src/Problem.cpp:143) */
  double gmres_seconds;     /* "Time to run the GMRES" (src/Problem.cpp:144) */
  int iterations;
  double estimate;
  double true_residual;
  long long dofs;
  int Gx, Gy;
} lsw_problem_result;

/* Number of problems in a config file (Input::getNumProblems). */
int lsw_config_num_problems(const char* path, int* n);
/* Parsed parameters of problem `id`; src_xy receives up to src_cap sources. */
int lsw_config_problem(const char* path, int id, lsw_problem_spec* spec, double* src_xy, int src_cap);
/* Problem::read + Problem::solve on the GPU.  x (optional, x_cap complex
 * entries) receives the solution on the global grid. */
int lsw_problem_run(const char* path, int id, int verbose, double* x, long long x_cap, lsw_problem_result* res);

/* ---- Output (OutputVTK::print2D, src/OutputVTK.cpp:82-366) ---------------
 * Writes <path>.vtk (print2D's legacy-ASCII layout of a one-rank run: per
 * cell its printed elements and their points, boundary points included) and
 * <path>.lswgrid (the same field on the global grid in full precision:
 * "LSWGRID1", int32 Gx, Gy, Gx x-coordinates, Gy y-coordinates, Gx*Gy
 * complex doubles).  x: global grid vector on the host.  Works on a
 * geometry-only DD. */
int lsw_dd_write_vtk(lsw_dd* dd, const double* x, const char* path);

/* ---- Serving ------------------------------------------------------------
 * No counterpart in the reference, whose driver reads, factorises and solves
 * one problem at a time (src/Problem.cpp:70-158).  A server keeps factorised
 * problems on the device (cache keyed by the registered problem = model and
 * omega), admits a problem when its estimated device bytes fit budget_bytes
 * (evicting idle problems, least recently used first), runs every source of
 * every request as one right-hand side of flexible GMRES, and batches the
 * right-hand sides of a problem per Arnoldi step: they join free slots between
 * steps and leave when their true relative residual is <= tol.  `workers`
 * host threads step different problems at once on their own streams.
 * submit/wait may be called from any thread. */
typedef struct lsw_server lsw_server;

typedef struct lsw_server_options {
  double budget_bytes; /* device bytes the cache may reserve */
  int slots;           /* right-hand sides in flight per problem (1..64) */
  int restart;         /* GMRES restart length (1..32) */
  double tol;          /* true relative residual target */
  int max_iter;        /* Arnoldi steps per right-hand side */
  int precision;       /* LSW_PREC_FP32 (mixed) or LSW_PREC_FP64 */
  int workers;         /* host threads (1..16) */
  int order;           /* LSW_ORDER_OLDEST: the problem with the oldest waiting
                        * request first; LSW_ORDER_BATCH (a batch submitted at
                        * once): problems already in flight first, then the
                        * one with the most estimated work (longest job first) */
} lsw_server_options;
#define LSW_ORDER_OLDEST 0
#define LSW_ORDER_BATCH 1

typedef struct lsw_request_result {
  int status;          /* LSW_OK or the error that failed the request */
  int problem;
  int nsrc;
  int max_iterations;  /* over the request's right-hand sides */
  double max_true_residual;
  double submit_s, start_s, end_s; /* seconds since lsw_server_create */
  char message[160];
} lsw_request_result;

typedef struct lsw_server_stats {
  long long requests, completed, solves, loads, evictions, steps;
  double reserved_bytes, peak_reserved_bytes;
  int resident;
  long long merged_steps; /* steps whose right-hand sides came from >= 2 requests */
  int max_active;         /* most right-hand sides in flight in one step */
  /* worker seconds summed over the workers: problem set-up (trees, schedule,
   * plans, assembly), factorisation, block plans + Krylov state, solve steps */
  double setup_s, factor_s, block_s, step_s;
} lsw_server_stats;

int lsw_server_create(const lsw_server_options* opt, lsw_server** out);
/* Stops the workers after their current step; pending requests are dropped. */
int lsw_server_destroy(lsw_server* s);
/* Registers a problem (parameters and velocity grid copied; precision taken
 * from the options); Gx, Gy: global grid of its solutions. */
int lsw_server_add_problem(lsw_server* s, const lsw_dd_params* p, int* problem, int* Gx, int* Gy);
/* Estimated device bytes of a registered problem: steady (factors, sweep and
 * Krylov buffers) and peak (while factorising); admission reserves these. */
int lsw_server_problem_bytes(lsw_server* s, int problem, double* steady, double* peak);
/* Queues nsrc point sources (src_xy: x0, y0, x1, y1, ...); x (optional, host)
 * receives nsrc global vectors and must stay valid until the wait returns. */
int lsw_server_submit(lsw_server* s, int problem, int nsrc, const double* src_xy, double* x, long long* request);
/* Blocks until the request is done (each request id is waited once). */
int lsw_server_wait(lsw_server* s, long long request, lsw_request_result* res);
int lsw_server_get_stats(lsw_server* s, lsw_server_stats* st);

#ifdef __cplusplus
}
#endif

#endif /* LSWEEPS_HIP_H */
