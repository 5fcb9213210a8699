// GlobalIterativeSolverGMRESHIP (see lsw_gmres.h).
#include "lsw_gmres.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "lsw_cplx.h"

namespace lsw {

namespace {
constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

// sum conj(a_i) b_i per block (fixed order within and across blocks)
__global__ __launch_bounds__(kRedThreads) void k_dotc_part(const double2* __restrict__ a,
                                                           const double2* __restrict__ b, long long n,
                                                           double2* __restrict__ part) {
  __shared__ double2 sm[kRedThreads];
  double2 acc = make_double2(0.0, 0.0);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc = cfma_conj(a[i], b[i], acc);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = cadd(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(kRedThreads) void k_norm_part(const double2* __restrict__ a, long long n,
                                                           double2* __restrict__ part) {
  __shared__ double sm[kRedThreads];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc = fma(a[i].x, a[i].x, fma(a[i].y, a[i].y, acc));
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sm[0], 0.0);
}

__global__ __launch_bounds__(kRedThreads) void k_final(const double2* __restrict__ part, int np,
                                                       double2* __restrict__ out) {
  __shared__ double2 sm[kRedThreads];
  double2 acc = make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc = cadd(acc, part[i]);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = cadd(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// y -= c * x with c read from device memory (optionally conjugated)
__global__ void k_axpy_dev(double2* __restrict__ y, const double2* __restrict__ x,
                           const double2* __restrict__ c, int conjc, long long n) {
  double2 cc = *c;
  if (conjc) cc.y = -cc.y;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = cfms(cc, x[i], y[i]);
}

// y = x / sqrt(nrm2sq)
__global__ void k_scale_invnorm(double2* __restrict__ y, const double2* __restrict__ x,
                                const double2* __restrict__ nrm2, long long n) {
  const double inv = 1.0 / sqrt(nrm2->x);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = make_double2(x[i].x * inv, x[i].y * inv);
}

// y += c x (host scalar)
__global__ void k_axpy(double2* __restrict__ y, const double2* __restrict__ x, double2 c, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = cfma(c, x[i], y[i]);
}

// y = a - b
__global__ void k_sub(double2* __restrict__ y, const double2* __restrict__ a, const double2* __restrict__ b,
                      long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = csub(a[i], b[i]);
}

// y = c x (host scalar)
__global__ void k_scale(double2* __restrict__ y, const double2* __restrict__ x, double2 c, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = cmul(c, x[i]);
}

inline dim3 gridFor(long long n) { return dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)); }
}  // namespace

namespace blas {
void dotc(const double2* a, const double2* b, long long n, double2* part, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dotc_part, dim3(kRedBlocks), dim3(kRedThreads), 0, s, a, b, n, part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(kRedThreads), 0, s, part, kRedBlocks, out);
}
void norm2sq(const double2* a, long long n, double2* part, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_norm_part, dim3(kRedBlocks), dim3(kRedThreads), 0, s, a, n, part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(kRedThreads), 0, s, part, kRedBlocks, out);
}
}  // namespace blas

GlobalIterativeSolverGMRESHIP::~GlobalIterativeSolverGMRESHIP() {
  for (double2* p : {dV, dZ, dT, dPart, dScal, dBV, dBZ, dBT, dBPart, dBH}) if (p) (void)devFree(p);
  if (dBLive) (void)devFree(dBLive);
}

GlobalIterativeSolver* GlobalIterativeSolverGMRESHIP::copy() const {
  auto* g = new GlobalIterativeSolverGMRESHIP();
  g->mParent = mParent;
  g->mTol = mTol;
  g->mMaxIter = mMaxIter;
  g->mRestart = mRestart;
  g->mMode = mMode;
  g->mVerbose = mVerbose;
  return g;
}

void GlobalIterativeSolverGMRESHIP::ensure(long long n, int vecs) const {
  LSW_REQUIRE(vecs + 1 <= 128, LSW_ERR_ARG, "GMRES: MaxIter/restart must be at most 126");
  if (n == mN && vecs <= mVecs && (mMode != kFlexible || dZ != nullptr)) return;
  for (double2* p : {dV, dZ, dT}) if (p) (void)devFree(p);
  dV = dZ = dT = nullptr;
  LSW_HIP(devMalloc(&dV, (size_t)vecs * n * sizeof(double2)));
  if (mMode == kFlexible) LSW_HIP(devMalloc(&dZ, (size_t)vecs * n * sizeof(double2)));
  LSW_HIP(devMalloc(&dT, (size_t)3 * n * sizeof(double2)));
  if (!dPart) LSW_HIP(devMalloc(&dPart, kRedBlocks * sizeof(double2)));
  if (!dScal) LSW_HIP(devMalloc(&dScal, 128 * sizeof(double2)));
  mN = n;
  mVecs = vecs;
}

namespace {
double2 toD2(Complex c) { return make_double2(c.real(), c.imag()); }
Complex fromD2(double2 c) { return Complex(c.x, c.y); }

double devNorm(const double2* a, long long n, double2* part, double2* scal, hipStream_t s) {
  blas::norm2sq(a, n, part, scal, s);
  double2 h;
  LSW_HIP(hipMemcpyAsync(&h, scal, sizeof(double2), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipStreamSynchronize(s));
  return std::sqrt(h.x);
}
}  // namespace

// src/GlobalIterativeSolverGMRES.cpp:65-348
void GlobalIterativeSolverGMRESHIP::solveReference(const DDCheckerboard& dd, const double2* b, double2* x,
                                                   GmresInfo& info, hipStream_t s) const {
  const long long n = dd.N();
  ensure(n, mMaxIter + 1);
  double2* rp = dT;          // M^{-1} b
  double2* t1 = dT + n;
  double2* t2 = dT + 2 * n;
  dd.applyPrecondDev(b, rp, s);
  info.preconditionerApplies++;
  LSW_HIP(hipMemcpyAsync(x, rp, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
  dd.applyADev(x, t1, s);
  info.matvecs++;
  dd.applyPrecondDev(t1, t2, s);
  info.preconditionerApplies++;
  double2* r = dV;  // Q[0] will be r / ||r||
  hipLaunchKernelGGL(k_sub, gridFor(n), dim3(256), 0, s, r, rp, t2, n);
  const double rhsNorm = devNorm(rp, n, dPart, dScal, s);
  const double rNorm = devNorm(r, n, dPart, dScal, s);
  info.log.assign(1, rNorm / rhsNorm);
  if (mVerbose) std::printf("iteration 0: %g\n", info.log[0]);
  info.estimate = info.log[0];
  if (info.log[0] < mTol || mMaxIter == 0) return;
  hipLaunchKernelGGL(k_scale, gridFor(n), dim3(256), 0, s, dV, r, make_double2(1.0 / rNorm, 0.0), n);
  std::vector<Complex> beta(1, Complex(rNorm, 0.0)), cs, sn;
  std::vector<std::vector<Complex>> Hcols;
  int lastk = 0;
  std::vector<double2> hh(mMaxIter + 2);
  for (int k = 0; k < mMaxIter; ++k) {
    double2* qk = dV + (long long)k * n;
    double2* w = dV + (long long)(k + 1) * n;
    dd.applyADev(qk, t1, s);
    info.matvecs++;
    dd.applyPrecondDev(t1, w, s);
    info.preconditionerApplies++;
    // modified Gram-Schmidt with h_i = sum conj(w) q_i (Eigen w.dot(q_i))
    for (int i = 0; i <= k; ++i) {
      blas::dotc(w, dV + (long long)i * n, n, dPart, dScal + i, s);
      hipLaunchKernelGGL(k_axpy_dev, gridFor(n), dim3(256), 0, s, w, dV + (long long)i * n, dScal + i, 0, n);
    }
    blas::norm2sq(w, n, dPart, dScal + k + 1, s);
    hipLaunchKernelGGL(k_scale_invnorm, gridFor(n), dim3(256), 0, s, w, w, dScal + k + 1, n);
    LSW_HIP(hipMemcpyAsync(hh.data(), dScal, (k + 2) * sizeof(double2), hipMemcpyDeviceToHost, s));
    LSW_HIP(hipStreamSynchronize(s));
    std::vector<Complex> h(k + 2);
    for (int i = 0; i <= k; ++i) h[i] = fromD2(hh[i]);
    h[k + 1] = Complex(std::sqrt(hh[k + 1].x), 0.0);
    LSW_REQUIRE(std::isfinite(h[k + 1].real()), LSW_ERR_NUMERIC, "GMRES breakdown: non-finite Arnoldi norm");
    // apply_givens_rotation: rotations 0..k-2 only (:315-320)
    for (int i = 1; i < k; ++i) {
      const Complex tmp = cs[i - 1] * h[i - 1] + sn[i - 1] * h[i];
      h[i] = -sn[i - 1] * h[i - 1] + cs[i - 1] * h[i];
      h[i - 1] = tmp;
    }
    // givens_rotation (:333-348)
    Complex c, sv;
    if (h[k] == 0.0) {
      c = 0.0;
      sv = 1.0;
    } else {
      const Complex tt = std::pow(std::pow(h[k], 2.0) + std::pow(h[k + 1], 2.0), 0.5);
      c = std::abs(h[k]) / tt;
      sv = c * h[k + 1] / h[k];
    }
    h[k] = c * h[k] + sv * h[k + 1];
    h[k + 1] = 0.0;
    cs.push_back(c);
    sn.push_back(sv);
    Hcols.push_back(h);
    beta.push_back(-sn[k] * beta[k]);
    beta[k] *= cs[k];
    info.log.push_back(std::abs(beta.back()) / rhsNorm);
    info.iterations = k + 1;
    if (mVerbose) std::printf("iteration %d: %g\n", k + 1, info.log.back());
    if (info.log.back() <= mTol || k == mMaxIter - 1) {
      lastk = k;
      break;
    }
  }
  info.estimate = info.log.back();
  // y from the leading (k+1)x(k+1) block (upper triangular after rotation)
  const int k = lastk;
  std::vector<Complex> y(k + 1);
  for (int i = k; i >= 0; --i) {
    Complex acc = beta[i];
    for (int j = i + 1; j <= k; ++j) acc -= Hcols[j][i] * y[j];
    y[i] = acc / Hcols[i][i];
  }
  for (int i = 0; i <= k; ++i)
    hipLaunchKernelGGL(k_axpy, gridFor(n), dim3(256), 0, s, x, dV + (long long)i * n, toD2(y[i]), n);
}

// Flexible GMRES(m), right preconditioning.
void GlobalIterativeSolverGMRESHIP::solveFlexible(const DDCheckerboard& dd, const double2* b, double2* x,
                                                  GmresInfo& info, hipStream_t s) const {
  const long long n = dd.N();
  const int m = std::max(1, mRestart);
  ensure(n, m + 1);
  double2* r = dT;
  double2* t1 = dT + n;
  LSW_HIP(hipMemsetAsync(x, 0, n * sizeof(double2), s));
  const double bn = devNorm(b, n, dPart, dScal, s);
  info.log.clear();
  if (bn == 0.0) {
    info.log.push_back(0.0);
    info.trueResidual = 0.0;
    return;
  }
  int it = 0;
  bool first = true;
  while (it < mMaxIter) {
    double beta;
    if (first) {
      LSW_HIP(hipMemcpyAsync(r, b, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
      beta = bn;
      first = false;
    } else {
      dd.applyADev(x, t1, s);
      info.matvecs++;
      hipLaunchKernelGGL(k_sub, gridFor(n), dim3(256), 0, s, r, b, t1, n);
      beta = devNorm(r, n, dPart, dScal, s);
    }
    info.log.push_back(beta / bn);
    if (beta / bn <= mTol) break;
    hipLaunchKernelGGL(k_scale, gridFor(n), dim3(256), 0, s, dV, r, make_double2(1.0 / beta, 0.0), n);
    std::vector<std::vector<Complex>> H(m + 1, std::vector<Complex>(m, 0.0));
    std::vector<Complex> cs(m), sn(m), g(m + 1, 0.0);
    g[0] = beta;
    int jend = 0;
    std::vector<double2> hh(m + 2);
    for (int j = 0; j < m && it < mMaxIter; ++j) {
      double2* vj = dV + (long long)j * n;
      double2* zj = dZ + (long long)j * n;
      double2* w = dV + (long long)(j + 1) * n;
      dd.applyPrecondDev(vj, zj, s);
      info.preconditionerApplies++;
      dd.applyADev(zj, w, s);
      info.matvecs++;
      for (int i = 0; i <= j; ++i) {
        blas::dotc(dV + (long long)i * n, w, n, dPart, dScal + i, s);  // q_i^H w
        hipLaunchKernelGGL(k_axpy_dev, gridFor(n), dim3(256), 0, s, w, dV + (long long)i * n, dScal + i, 0, n);
      }
      blas::norm2sq(w, n, dPart, dScal + j + 1, s);
      hipLaunchKernelGGL(k_scale_invnorm, gridFor(n), dim3(256), 0, s, w, w, dScal + j + 1, n);
      LSW_HIP(hipMemcpyAsync(hh.data(), dScal, (j + 2) * sizeof(double2), hipMemcpyDeviceToHost, s));
      LSW_HIP(hipStreamSynchronize(s));
      for (int i = 0; i <= j; ++i) H[i][j] = fromD2(hh[i]);
      H[j + 1][j] = std::sqrt(hh[j + 1].x);
      LSW_REQUIRE(std::isfinite(H[j + 1][j].real()), LSW_ERR_NUMERIC, "GMRES breakdown: non-finite Arnoldi norm");
      for (int i = 0; i < j; ++i) {
        const Complex tmp = std::conj(cs[i]) * H[i][j] + std::conj(sn[i]) * H[i + 1][j];
        H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
        H[i][j] = tmp;
      }
      const Complex a = H[j][j], bb = H[j + 1][j];
      const double den = std::sqrt(std::norm(a) + std::norm(bb));
      cs[j] = a / den;
      sn[j] = bb / den;
      H[j][j] = std::conj(cs[j]) * a + std::conj(sn[j]) * bb;
      H[j + 1][j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = std::conj(cs[j]) * g[j];
      ++it;
      jend = j + 1;
      info.log.push_back(std::abs(g[j + 1]) / bn);
      if (mVerbose) std::printf("iteration %d: %g\n", it, info.log.back());
      if (std::abs(g[j + 1]) / bn <= mTol) break;
    }
    std::vector<Complex> y(jend);
    for (int i = jend - 1; i >= 0; --i) {
      Complex acc = g[i];
      for (int k2 = i + 1; k2 < jend; ++k2) acc -= H[i][k2] * y[k2];
      y[i] = acc / H[i][i];
    }
    for (int i = 0; i < jend; ++i)
      hipLaunchKernelGGL(k_axpy, gridFor(n), dim3(256), 0, s, x, dZ + (long long)i * n, toD2(y[i]), n);
    if (info.log.back() <= mTol) {
      // confirm with the true residual before stopping
      dd.applyADev(x, t1, s);
      info.matvecs++;
      hipLaunchKernelGGL(k_sub, gridFor(n), dim3(256), 0, s, r, b, t1, n);
      const double tr = devNorm(r, n, dPart, dScal, s) / bn;
      if (tr <= mTol) break;
    }
  }
  info.iterations = it;
  info.estimate = info.log.back();
}

void GlobalIterativeSolverGMRESHIP::solveDev(const double2* b, double2* x, GmresInfo* infoOut, hipStream_t s) const {
  const auto* dd = dynamic_cast<const DDCheckerboard*>(mParent);
  LSW_REQUIRE(dd != nullptr, LSW_ERR_STATE, "GMRESHIP needs a DDCheckerboard parent");
  if (!s) s = dd->stream();
  GmresInfo info;
  if (mMode == kReference)
    solveReference(*dd, b, x, info, s);
  else
    solveFlexible(*dd, b, x, info, s);
  // true FP64 residual
  const long long n = dd->N();
  double2* t1 = dT + n;
  double2* r = dT;
  dd->applyADev(x, t1, s);
  hipLaunchKernelGGL(k_sub, gridFor(n), dim3(256), 0, s, r, b, t1, n);
  const double rn = devNorm(r, n, dPart, dScal, s);
  const double bn = devNorm(b, n, dPart, dScal, s);
  info.trueResidual = bn > 0 ? rn / bn : rn;
  LSW_HIP(hipGetLastError());
  mInfo = info;
  if (infoOut) *infoOut = info;
}

void GlobalIterativeSolverGMRESHIP::solve(const std::vector<ComplexVector>& rhs, std::vector<ComplexVector>& res) const {
  const auto* dd = dynamic_cast<const DDCheckerboard*>(mParent);
  LSW_REQUIRE(dd != nullptr, LSW_ERR_STATE, "GMRESHIP needs a DDCheckerboard parent");
  std::vector<Complex> g;
  dd->scatterOwned(rhs, g);
  const long long n = dd->N();
  double2 *db = nullptr, *dx = nullptr;
  LSW_HIP(devMalloc(&db, n * sizeof(double2)));
  LSW_HIP(devMalloc(&dx, n * sizeof(double2)));
  LSW_HIP(uploadSync(db, g.data(), n * sizeof(double2)));
  try {
    solveDev(db, dx, nullptr, dd->stream());
    LSW_HIP(hipStreamSynchronize(dd->stream()));
    LSW_HIP(hipMemcpy(g.data(), dx, n * sizeof(double2), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)devFree(db);
    (void)devFree(dx);
    throw;
  }
  (void)devFree(db);
  (void)devFree(dx);
  dd->gatherOwned(g, res);
}

}  // namespace lsw
