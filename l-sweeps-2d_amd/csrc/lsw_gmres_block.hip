// Block flexible GMRES (see lsw_gmres.h, solveBlockDev).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "lsw_cplx.h"
#include "lsw_gmres.h"

namespace lsw {

namespace {
constexpr int kBlocks = 256;
constexpr int kThreads = 256;

// Column a of the step's vectors for grid row y: live[y] while some of the
// active right-hand sides have converged (the inner steps then run only on
// the others), y itself otherwise.
__device__ __forceinline__ int liveCol(const int* __restrict__ live, int y) { return live ? live[y] : y; }

// part[(pair)*kBlocks + block] = sum conj(V_i^a) w^a over the block's range;
// pair = y*nv + i with a = liveCol(y), V_i^a = V + (i*nact + a)*n, w^a = W + a*n.
__global__ __launch_bounds__(kThreads) void k_bdot_part(const double2* __restrict__ V, const double2* __restrict__ W,
                                                        int nv, int nact, long long n, double2* __restrict__ part,
                                                        const int* __restrict__ live) {
  __shared__ double2 sm[kThreads];
  const int pair = blockIdx.y;
  const int y = pair / nv, i = pair - y * nv;
  const int a = liveCol(live, y);
  const double2* v = V + ((long long)i * nact + a) * n;
  const double2* w = W + (long long)a * n;
  double2 acc = make_double2(0.0, 0.0);
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    acc = cfma_conj(v[q], w[q], acc);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = cadd(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)pair * kBlocks + blockIdx.x] = sm[0];
}

// per-vector squared norms: part[y*kBlocks + block] of column liveCol(y)
__global__ __launch_bounds__(kThreads) void k_bnorm_part(const double2* __restrict__ W, long long n,
                                                         double2* __restrict__ part,
                                                         const int* __restrict__ live = nullptr) {
  __shared__ double sm[kThreads];
  const double2* w = W + (long long)liveCol(live, blockIdx.y) * n;
  double acc = 0.0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    acc = fma(w[q].x, w[q].x, fma(w[q].y, w[q].y, acc));
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)blockIdx.y * kBlocks + blockIdx.x] = make_double2(sm[0], 0.0);
}

// out[p] = sum_b part[p*kBlocks + b] (+= when accumulate)
__global__ __launch_bounds__(kThreads) void k_bfinal(const double2* __restrict__ part, double2* __restrict__ out,
                                                     int accumulate) {
  __shared__ double2 sm[kThreads];
  const int p = blockIdx.x;
  double2 acc = make_double2(0.0, 0.0);
  for (int b = threadIdx.x; b < kBlocks; b += blockDim.x) acc = cadd(acc, part[(long long)p * kBlocks + b]);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = cadd(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[p] = accumulate ? cadd(out[p], sm[0]) : sm[0];
}

// w^a -= sum_i h[y*nv+i] V_i^a, a = liveCol(y)
__global__ void k_bupdate(const double2* __restrict__ V, double2* __restrict__ W, const double2* __restrict__ h,
                          int nv, int nact, long long n, const int* __restrict__ live) {
  const int y = blockIdx.y, a = liveCol(live, y);
  double2* w = W + (long long)a * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    double2 acc = w[q];
    for (int i = 0; i < nv; ++i) acc = cfms(h[y * nv + i], V[((long long)i * nact + a) * n + q], acc);
    w[q] = acc;
  }
}

// w^a *= 1/sqrt(nrm2[y]), a = liveCol(y)
__global__ void k_bscale(double2* __restrict__ W, const double2* __restrict__ nrm2, long long n,
                         const int* __restrict__ live) {
  const int y = blockIdx.y;
  const double inv = 1.0 / sqrt(nrm2[y].x);
  double2* w = W + (long long)liveCol(live, y) * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    w[q] = make_double2(w[q].x * inv, w[q].y * inv);
}

// gather (dst column y = src column live[y]) or scatter (dst column live[y] =
// src column y) of n-vectors
__global__ void k_cols_copy(double2* __restrict__ dst, const double2* __restrict__ src, const int* __restrict__ live,
                            long long n, int scatter) {
  const long long c = live[blockIdx.y];
  double2* d = dst + (scatter ? c : (long long)blockIdx.y) * n;
  const double2* v = src + (scatter ? (long long)blockIdx.y : c) * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    d[q] = v[q];
}

__global__ void k_gather_scale(double2* __restrict__ dst, const double2* __restrict__ src, double c, long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = make_double2(src[q].x * c, src[q].y * c);
}

__global__ void k_axpy_h(double2* __restrict__ y, const double2* __restrict__ x, double2 c, long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    y[q] = cfma(c, x[q], y[q]);
}

__global__ void k_sub2(double2* __restrict__ y, const double2* __restrict__ a, const double2* __restrict__ b,
                       long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    y[q] = csub(a[q], b[q]);
}

inline dim3 gridN(long long n) { return dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)); }
}  // namespace

void GlobalIterativeSolverGMRESHIP::solveBlockDev(const double2* B, double2* X, int nrhs,
                                                  std::vector<GmresInfo>* infosOut, hipStream_t s) const {
  const auto* dd = dynamic_cast<const DDCheckerboard*>(mParent);
  LSW_REQUIRE(dd != nullptr, LSW_ERR_STATE, "GMRESHIP needs a DDCheckerboard parent");
  LSW_REQUIRE(nrhs >= 1, LSW_ERR_ARG, "solveBlock: nrhs must be positive");
  if (!s) s = dd->stream();
  const long long n = dd->N();
  const int m = std::max(1, mRestart);
  if (n != mBN || nrhs > mBRhs || m > mBM) {
    for (double2* p : {dBV, dBZ, dBT, dBPart, dBH}) if (p) (void)devFree(p);
    if (dBLive) (void)devFree(dBLive);
    dBV = dBZ = dBT = dBPart = dBH = nullptr;
    dBLive = nullptr;
    LSW_HIP(devMalloc(&dBV, (size_t)(m + 1) * nrhs * n * sizeof(double2)));
    LSW_HIP(devMalloc(&dBZ, (size_t)m * nrhs * n * sizeof(double2)));
    LSW_HIP(devMalloc(&dBT, (size_t)2 * nrhs * n * sizeof(double2)));
    LSW_HIP(devMalloc(&dBPart, (size_t)(m + 1) * nrhs * kBlocks * sizeof(double2)));
    LSW_HIP(devMalloc(&dBH, (size_t)2 * (m + 2) * nrhs * sizeof(double2)));
    LSW_HIP(devMalloc(&dBLive, (size_t)nrhs * sizeof(int)));
    mBN = n;
    mBRhs = nrhs;
    mBM = m;
  }
  double2* R = dBT;            // residuals of all right-hand sides
  double2* T = dBT + nrhs * n;  // A x / A z
  std::vector<GmresInfo> infos(nrhs);
  // ||b_r||
  std::vector<double> bn(nrhs), res(nrhs);
  auto norms = [&](const double2* W, int k, std::vector<double>& out) {
    hipLaunchKernelGGL(k_bnorm_part, dim3(kBlocks, k), dim3(kThreads), 0, s, W, n, dBPart);
    hipLaunchKernelGGL(k_bfinal, dim3(k), dim3(kThreads), 0, s, dBPart, dBH, 0);
    std::vector<double2> h(k);
    LSW_HIP(hipMemcpyAsync(h.data(), dBH, k * sizeof(double2), hipMemcpyDeviceToHost, s));
    LSW_HIP(hipStreamSynchronize(s));
    out.resize(k);
    for (int a = 0; a < k; ++a) out[a] = std::sqrt(h[a].x);
  };
  norms(B, nrhs, bn);
  LSW_HIP(hipMemsetAsync(X, 0, (size_t)nrhs * n * sizeof(double2), s));
  LSW_HIP(hipMemcpyAsync(R, B, (size_t)nrhs * n * sizeof(double2), hipMemcpyDeviceToDevice, s));
  res = bn;
  // iters[r] counts the Arnoldi steps of right-hand side r.  A right-hand side
  // is active only while iters[r] < mMaxIter, every inner step raises iters of
  // each active one that is not done, and an outer pass runs at least one
  // inner step, so the loop ends after at most nrhs * mMaxIter steps.
  std::vector<int> iters(nrhs, 0);
  bool first = true;
  bool fresh = false;  // res holds the true residuals of the final X
  while (true) {
    if (!first) {
      dd->applyADevBlock(X, T, nrhs, s);
      hipLaunchKernelGGL(k_sub2, gridN(nrhs * n), dim3(256), 0, s, R, B, T, nrhs * n);
      norms(R, nrhs, res);
    }
    first = false;
    std::vector<int> act;
    for (int r = 0; r < nrhs; ++r) {
      const double rel = bn[r] > 0 ? res[r] / bn[r] : 0.0;
      infos[r].log.push_back(rel);
      infos[r].estimate = rel;
      if (rel > mTol && iters[r] < mMaxIter) act.push_back(r);
    }
    if (act.empty()) {
      fresh = true;
      break;
    }
    const int na = (int)act.size();
    // V_0 = r / ||r|| for the active right-hand sides (compacted)
    for (int a = 0; a < na; ++a)
      hipLaunchKernelGGL(k_gather_scale, gridN(n), dim3(256), 0, s, dBV + (long long)a * n, R + (long long)act[a] * n,
                         1.0 / res[act[a]], n);
    std::vector<std::vector<std::vector<Complex>>> H(na, std::vector<std::vector<Complex>>(m + 1, std::vector<Complex>(m, 0.0)));
    std::vector<std::vector<Complex>> cs(na, std::vector<Complex>(m)), sn(na, std::vector<Complex>(m)),
        g(na, std::vector<Complex>(m + 1, 0.0));
    std::vector<int> jend(na, 0);
    std::vector<bool> done(na, false);
    for (int a = 0; a < na; ++a) g[a][0] = res[act[a]];
    int j = 0;
    for (; j < m; ++j) {
      double2* Vj = dBV + (long long)j * na * n;
      double2* Zj = dBZ + (long long)j * na * n;
      double2* W = dBV + (long long)(j + 1) * na * n;
      // the step runs on the right-hand sides whose estimate has not
      // converged: once some have, the others are gathered into a block of
      // the next power of two (zero-padded; a handful of plan widths) for
      // M^-1 and A and scattered back, and the Gram-Schmidt kernels index
      // them through the live list
      std::vector<int> live;
      for (int a = 0; a < na; ++a)
        if (!done[a]) live.push_back(a);
      const int nl = (int)live.size();
      const int* dl = nullptr;
      if (nl < na) {
        LSW_HIP(hipMemcpyAsync(dBLive, live.data(), nl * sizeof(int), hipMemcpyHostToDevice, s));
        dl = dBLive;
        int wb = 1;
        while (wb < nl) wb *= 2;
        wb = std::min(wb, na);
        double2* G = dBT;                         // gathered V_j, then A M^-1 V_j
        double2* GZ = dBT + (long long)nrhs * n;  // M^-1 V_j
        const dim3 cg(gridN(n).x, nl);
        hipLaunchKernelGGL(k_cols_copy, cg, dim3(256), 0, s, G, Vj, dl, n, 0);
        if (wb > nl) LSW_HIP(hipMemsetAsync(G + (long long)nl * n, 0, (size_t)(wb - nl) * n * sizeof(double2), s));
        dd->applyPrecondDev(G, GZ, s, wb);
        hipLaunchKernelGGL(k_cols_copy, cg, dim3(256), 0, s, Zj, GZ, dl, n, 1);
        dd->applyADevBlock(GZ, G, nl, s);
        hipLaunchKernelGGL(k_cols_copy, cg, dim3(256), 0, s, W, G, dl, n, 1);
      } else {
        dd->applyPrecondDev(Vj, Zj, s, na);
        dd->applyADevBlock(Zj, W, na, s);
      }
      for (int a : live) infos[act[a]].preconditionerApplies++, infos[act[a]].matvecs++;
      const int nv = j + 1;
      // two passes of classical Gram-Schmidt
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_bdot_part, dim3(kBlocks, nv * nl), dim3(kThreads), 0, s, dBV, W, nv, na, n, dBPart, dl);
        hipLaunchKernelGGL(k_bfinal, dim3(nv * nl), dim3(kThreads), 0, s, dBPart, dBH + pass * (m + 2) * nrhs, 0);
        hipLaunchKernelGGL(k_bupdate, dim3(gridN(n).x, nl), dim3(256), 0, s, dBV, W, dBH + pass * (m + 2) * nrhs,
                           nv, na, n, dl);
      }
      hipLaunchKernelGGL(k_bnorm_part, dim3(kBlocks, nl), dim3(kThreads), 0, s, W, n, dBPart, dl);
      hipLaunchKernelGGL(k_bfinal, dim3(nl), dim3(kThreads), 0, s, dBPart, dBH + 2 * (m + 2) * nrhs - nl, 0);
      hipLaunchKernelGGL(k_bscale, dim3(gridN(n).x, nl), dim3(256), 0, s, W, dBH + 2 * (m + 2) * nrhs - nl, n, dl);
      std::vector<double2> h1(nv * nl), h2(nv * nl), nr(nl);
      LSW_HIP(hipMemcpyAsync(h1.data(), dBH, nv * nl * sizeof(double2), hipMemcpyDeviceToHost, s));
      LSW_HIP(hipMemcpyAsync(h2.data(), dBH + (m + 2) * nrhs, nv * nl * sizeof(double2), hipMemcpyDeviceToHost, s));
      LSW_HIP(hipMemcpyAsync(nr.data(), dBH + 2 * (m + 2) * nrhs - nl, nl * sizeof(double2), hipMemcpyDeviceToHost, s));
      LSW_HIP(hipStreamSynchronize(s));
      bool all = true;
      for (int y = 0; y < nl; ++y) {
        const int a = live[y];
        auto& Ha = H[a];
        for (int i = 0; i < nv; ++i)
          Ha[i][j] = Complex(h1[y * nv + i].x + h2[y * nv + i].x, h1[y * nv + i].y + h2[y * nv + i].y);
        Ha[j + 1][j] = std::sqrt(nr[y].x);
        LSW_REQUIRE(std::isfinite(Ha[j + 1][j].real()), LSW_ERR_NUMERIC, "GMRES breakdown: non-finite norm");
        for (int i = 0; i < j; ++i) {
          const Complex t = std::conj(cs[a][i]) * Ha[i][j] + std::conj(sn[a][i]) * Ha[i + 1][j];
          Ha[i + 1][j] = -sn[a][i] * Ha[i][j] + cs[a][i] * Ha[i + 1][j];
          Ha[i][j] = t;
        }
        const Complex x0 = Ha[j][j], x1 = Ha[j + 1][j];
        const double den = std::sqrt(std::norm(x0) + std::norm(x1));
        cs[a][j] = x0 / den;
        sn[a][j] = x1 / den;
        Ha[j][j] = std::conj(cs[a][j]) * x0 + std::conj(sn[a][j]) * x1;
        Ha[j + 1][j] = 0.0;
        g[a][j + 1] = -sn[a][j] * g[a][j];
        g[a][j] = std::conj(cs[a][j]) * g[a][j];
        jend[a] = j + 1;
        const int r = act[a];
        iters[r]++;
        const double est = std::abs(g[a][j + 1]) / bn[r];
        infos[r].log.push_back(est);
        if (mVerbose) std::printf("rhs %d iteration %d: %g\n", r, iters[r], est);
        if (est <= mTol || iters[r] >= mMaxIter) done[a] = true;
        else all = false;
      }
      if (all) {
        ++j;
        break;
      }
    }
    // x_r += Z y_r
    for (int a = 0; a < na; ++a) {
      const int je = jend[a];
      std::vector<Complex> y(je);
      for (int i = je - 1; i >= 0; --i) {
        Complex acc = g[a][i];
        for (int k = i + 1; k < je; ++k) acc -= H[a][i][k] * y[k];
        y[i] = acc / H[a][i][i];
      }
      for (int i = 0; i < je; ++i)
        hipLaunchKernelGGL(k_axpy_h, gridN(n), dim3(256), 0, s, X + (long long)act[a] * n,
                           dBZ + ((long long)i * na + a) * n, make_double2(y[i].real(), y[i].imag()), n);
    }
    bool anyLeft = false;
    for (int r = 0; r < nrhs; ++r) anyLeft |= iters[r] < mMaxIter;
    if (!anyLeft) break;
  }
  // true FP64 residuals (computed at the top of the last pass when the loop
  // ended there)
  if (!fresh) {
    dd->applyADevBlock(X, T, nrhs, s);
    hipLaunchKernelGGL(k_sub2, gridN(nrhs * n), dim3(256), 0, s, R, B, T, nrhs * n);
    norms(R, nrhs, res);
  }
  for (int r = 0; r < nrhs; ++r) {
    infos[r].iterations = iters[r];
    infos[r].trueResidual = bn[r] > 0 ? res[r] / bn[r] : res[r];
  }
  LSW_HIP(hipGetLastError());
  if (infosOut) *infosOut = infos;
}

}  // namespace lsw
