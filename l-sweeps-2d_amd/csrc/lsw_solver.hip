// LocalLinearSolverHIP (see lsw_solver.h).
#include "lsw_solver.h"

#include <algorithm>
#include <cstring>

namespace lsw {

LocalLinearSolverHIP::~LocalLinearSolverHIP() {
  for (void* p : {dX, dR, dG}) if (p) (void)devFree(p);
  if (dTasks) (void)devFree(dTasks);
  mFS.reset();
  if (mStream) (void)hipStreamDestroy(mStream);
}

// Recover the nx x ny grid of a 5-point matrix with index i*ny+j and check the
// whole pattern (LocalLinearSolverEigenSparseLU accepts any sparse matrix; the
// MI355X solver is specialised to the structured operators the assemblers
// produce).
void LocalLinearSolverHIP::setMatrix(ComplexSparseMatrix* A) {
  LSW_REQUIRE(A != nullptr, LSW_ERR_ARG, "setMatrix: null matrix");
  LSW_REQUIRE(A->rows == A->cols && A->rows > 0, LSW_ERR_ARG, "setMatrix: matrix must be square");
  LSW_REQUIRE((int)A->colptr.size() == A->cols + 1, LSW_ERR_ARG, "setMatrix: bad column pointer");
  const int N = A->rows;
  int ny = N;  // 1 x N grid unless column 0 couples to a far row
  for (int k = A->colptr[0]; k < A->colptr[1]; ++k) {
    const int r = A->rowind[k];
    if (r > 1) ny = std::min(ny == N ? r : ny, r);
  }
  LSW_REQUIRE(N % ny == 0, LSW_ERR_UNSUPPORTED, "matrix is not a 5-point structured-grid operator");
  const int nx = N / ny;
  for (int d = 0; d < N; ++d) {
    const int i = d / ny, j = d % ny;
    for (int k = A->colptr[d]; k < A->colptr[d + 1]; ++k) {
      const int q = A->rowind[k];
      LSW_REQUIRE(q >= 0 && q < N, LSW_ERR_ARG, "setMatrix: row index out of range");
      const int diff = d - q;
      const int qi = q / ny, qj = q % ny;
      const bool ok = diff == 0 || ((diff == 1 || diff == -1) && qi == i && std::abs(qj - j) == 1) ||
                      ((diff == ny || diff == -ny) && qj == j && ny < N);
      LSW_REQUIRE(ok, LSW_ERR_UNSUPPORTED, "matrix is not a 5-point structured-grid operator");
    }
  }
  mA = A;
  mNx = nx;
  mNy = ny;
  mFS.reset();
}

void LocalLinearSolverHIP::factorize() {
  LSW_REQUIRE(mA != nullptr, LSW_ERR_STATE, "factorize called before setMatrix");
  if (mFS && mFS->factorized()) return;  // idempotent (src/LocalLinearSolverEigenSparseLU.cpp:42-45)
  if (!mStream) LSW_HIP(hipStreamCreateWithFlags(&mStream, hipStreamNonBlocking));
  const int N = mA->rows, ny = mNy;
  std::vector<double2> coef((size_t)5 * N, make_double2(0.0, 0.0));
  for (int d = 0; d < N; ++d)
    for (int k = mA->colptr[d]; k < mA->colptr[d + 1]; ++k) {
      const int q = mA->rowind[k];
      const int diff = d - q;  // entry A[q][d]: coefficient of row q towards d
      const int dir = diff == 0 ? 0 : diff == -ny ? 1 : diff == ny ? 2 : diff == -1 ? 3 : 4;
      coef[(size_t)dir * N + q] = make_double2(mA->vals[k].real(), mA->vals[k].imag());
    }
  CellDev c;
  std::memset(&c, 0, sizeof(c));
  c.nx = mNx;
  c.ny = mNy;
  mFS = std::make_unique<FactorSet>();
  mFS->addCellWithStencil(c, coef);
  mFS->setup(mStream);
  mFS->factorize(mPrecision, mStream);
}

void LocalLinearSolverHIP::solveMany(const Complex* b, Complex* x, int nrhs) const {
  LSW_REQUIRE(mFS && mFS->factorized(), LSW_ERR_STATE, "Solve called before factorization");
  LSW_REQUIRE(nrhs >= 0, LSW_ERR_ARG, "negative right-hand-side count");
  if (nrhs == 0) return;
  const long long N = mA->rows;
  const long long G = std::max<long long>(mFS->maxG(), 1);
  const size_t es = mPrecision == LSW_FP64 ? sizeof(double2) : sizeof(float2);
  const int per = mFS->maxTaskRhs();
  const int ntasks = (nrhs + per - 1) / per;
  if (nrhs > mCap) {
    for (void* p : {dX, dR, dG}) if (p) (void)devFree(p);
    if (dTasks) (void)devFree(dTasks);
    // X, R and G are task-major: ntasks rows of up to per right-hand sides
    const size_t rows = (size_t)ntasks * per;
    LSW_HIP(devMalloc(&dX, rows * N * es));
    LSW_HIP(devMalloc(&dR, rows * N * es));
    LSW_HIP(devMalloc(&dG, rows * G * es));
    // the solve kernels read (and multiply by zero) entries of X, R and G
    // they do not use: they must hold finite values from the start
    LSW_HIP(hipMemsetAsync(dX, 0, rows * N * es, mStream));
    LSW_HIP(hipMemsetAsync(dR, 0, rows * N * es, mStream));
    LSW_HIP(hipMemsetAsync(dG, 0, rows * G * es, mStream));
    LSW_HIP(devMalloc(&dTasks, (size_t)ntasks * sizeof(SolveTask)));
    mCap = nrhs;
  }
  std::vector<SolveTask> tasks(ntasks);
  for (int t = 0; t < ntasks; ++t) {
    std::memset(&tasks[t], 0, sizeof(SolveTask));
    tasks[t].cell = 0;
    tasks[t].nrhs = std::min(per, nrhs - t * per);
    for (int r = 0; r < tasks[t].nrhs; ++r) tasks[t].slot[r] = t * per + r;
  }
  SolvePlan plan;
  mFS->buildPlan(tasks, {}, plan);
  LSW_REQUIRE(plan.taskRows() <= (long long)((mCap + per - 1) / per) * per, LSW_ERR_STATE,
              "solve scratch smaller than the plan needs");
  // right-hand side r = task r / per, lane r % per of the task-major buffer
  const int W = plan.width;
  const size_t cnt = (size_t)plan.taskRows() * N;
  auto pos = [&](int r, long long p) { return ((size_t)(r / per) * N + p) * W + (r % per); };
  if (mPrecision == LSW_FP64) {
    std::vector<double2> h(cnt, make_double2(0.0, 0.0));
    for (int r = 0; r < nrhs; ++r)
      for (long long p = 0; p < N; ++p) {
        const Complex v = b[(size_t)r * N + p];
        h[pos(r, p)] = make_double2(v.real(), v.imag());
      }
    LSW_HIP(hipMemcpyAsync(dX, h.data(), cnt * sizeof(double2), hipMemcpyHostToDevice, mStream));
    mFS->solve(plan, dX, dR, dG, N, G, mStream);
    LSW_HIP(hipMemcpyAsync(h.data(), dX, cnt * sizeof(double2), hipMemcpyDeviceToHost, mStream));
    LSW_HIP(hipStreamSynchronize(mStream));
    for (int r = 0; r < nrhs; ++r)
      for (long long p = 0; p < N; ++p) x[(size_t)r * N + p] = Complex(h[pos(r, p)].x, h[pos(r, p)].y);
  } else {
    std::vector<float2> h(cnt, make_float2(0.f, 0.f));
    for (int r = 0; r < nrhs; ++r)
      for (long long p = 0; p < N; ++p) {
        const Complex v = b[(size_t)r * N + p];
        h[pos(r, p)] = make_float2((float)v.real(), (float)v.imag());
      }
    LSW_HIP(hipMemcpyAsync(dX, h.data(), cnt * sizeof(float2), hipMemcpyHostToDevice, mStream));
    mFS->solve(plan, dX, dR, dG, N, G, mStream);
    LSW_HIP(hipMemcpyAsync(h.data(), dX, cnt * sizeof(float2), hipMemcpyDeviceToHost, mStream));
    LSW_HIP(hipStreamSynchronize(mStream));
    for (int r = 0; r < nrhs; ++r)
      for (long long p = 0; p < N; ++p) x[(size_t)r * N + p] = Complex(h[pos(r, p)].x, h[pos(r, p)].y);
  }
}

void LocalLinearSolverHIP::solve(ComplexVector& b, ComplexVector& res) const {
  LSW_REQUIRE(mA && (int)b.size() == mA->rows, LSW_ERR_ARG, "solve: right-hand side has the wrong length");
  res.resize(b.size());
  solveMany(b.data(), res.data(), 1);
}

}  // namespace lsw
