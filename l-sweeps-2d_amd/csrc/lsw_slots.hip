// GmresSlots: iteration-level batching of flexible GMRES (see lsw_slots.h).
#include "lsw_slots.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "lsw_cplx.h"
#include "lsw_gmres.h"

namespace lsw {

namespace {
constexpr int kBlocks = 256;
constexpr int kThreads = 256;
constexpr int kMaxRestart = 32;

// The active slots of one step, passed by value.  nv[a] = j + 1 basis vectors
// of slot slot[a] take part in its orthogonalisation.
struct SlotList {
  int n;
  int maxnv;
  int slot[kMaxSlots];
  int nv[kMaxSlots];
};

struct Weights {
  double2 y[kMaxRestart];
};

inline unsigned gridN(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 2048); }

// In[a] = V(slot[a], nv[a]-1)
__global__ void k_sgather(double2* __restrict__ In, const double2* __restrict__ V, SlotList L, int m, long long n) {
  const int a = blockIdx.y;
  const double2* src = V + ((long long)L.slot[a] * (m + 1) + L.nv[a] - 1) * n;
  double2* dst = In + (long long)a * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

// Z(slot[a], nv[a]-1) = PZ[a]
__global__ void k_sscatter(double2* __restrict__ Z, const double2* __restrict__ PZ, SlotList L, int m, long long n) {
  const int a = blockIdx.y;
  double2* dst = Z + ((long long)L.slot[a] * m + L.nv[a] - 1) * n;
  const double2* src = PZ + (long long)a * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

// part[pair*kBlocks + block] = sum conj(V(slot[a], i)) W[a], pair = a*maxnv + i
// (zero for i >= nv[a]); fixed reduction order.
__global__ __launch_bounds__(kThreads) void k_sdot_part(const double2* __restrict__ V, const double2* __restrict__ W,
                                                        SlotList L, int m, long long n, double2* __restrict__ part) {
  __shared__ double2 sm[kThreads];
  const int pair = blockIdx.y;
  const int a = pair / L.maxnv, i = pair - a * L.maxnv;
  if (i >= L.nv[a]) {
    if (threadIdx.x == 0) part[(long long)pair * kBlocks + blockIdx.x] = make_double2(0.0, 0.0);
    return;
  }
  const double2* v = V + ((long long)L.slot[a] * (m + 1) + i) * n;
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

__global__ __launch_bounds__(kThreads) void k_snorm_part(const double2* __restrict__ W, long long n,
                                                         double2* __restrict__ part) {
  __shared__ double sm[kThreads];
  const double2* w = W + (long long)blockIdx.y * n;
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

// out[p] = sum_b part[p*kBlocks + b]
__global__ __launch_bounds__(kThreads) void k_sfinal(const double2* __restrict__ part, double2* __restrict__ out) {
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
  if (threadIdx.x == 0) out[p] = sm[0];
}

// W[a] -= sum_{i < nv[a]} h[a*maxnv + i] V(slot[a], i)
__global__ void k_supdate(const double2* __restrict__ V, double2* __restrict__ W, const double2* __restrict__ h,
                          SlotList L, int m, long long n) {
  const int a = blockIdx.y;
  const int nv = L.nv[a];
  const double2* vb = V + (long long)L.slot[a] * (m + 1) * n;
  const double2* ha = h + a * L.maxnv;
  double2* w = W + (long long)a * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    double2 acc = w[q];
    for (int i = 0; i < nv; ++i) acc = cfms(ha[i], vb[(long long)i * n + q], acc);
    w[q] = acc;
  }
}

// V(slot[a], nv[a]) = W[a] / sqrt(nrm2[a])
__global__ void k_sscale_to(double2* __restrict__ V, const double2* __restrict__ W,
                            const double2* __restrict__ nrm2, SlotList L, int m, long long n) {
  const int a = blockIdx.y;
  const double inv = 1.0 / sqrt(nrm2[a].x);
  double2* dst = V + ((long long)L.slot[a] * (m + 1) + L.nv[a]) * n;
  const double2* src = W + (long long)a * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = make_double2(src[q].x * inv, src[q].y * inv);
}

// x += sum_{i < k} y_i Z_i (Z_i consecutive vectors of n entries)
__global__ void k_scomb(double2* __restrict__ x, const double2* __restrict__ Z, Weights w, int k, long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    double2 acc = x[q];
    for (int i = 0; i < k; ++i) acc = cfma(w.y[i], Z[(long long)i * n + q], acc);
    x[q] = acc;
  }
}

__global__ void k_ssub(double2* __restrict__ r, const double2* __restrict__ b, const double2* __restrict__ ax,
                       long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    r[q] = csub(b[q], ax[q]);
}

__global__ void k_sscale(double2* __restrict__ dst, const double2* __restrict__ src, double c, long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = make_double2(src[q].x * c, src[q].y * c);
}
}  // namespace

int planWidth(int na) {
  static const int w[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64};
  for (int x : w)
    if (x >= na) return x;
  return na;
}

size_t GmresSlots::bytesFor(long long n, int cap, int m) {
  const size_t vecs = (size_t)cap * (m + 1) + (size_t)cap * m + 3 * (size_t)cap + 2 * (size_t)planWidth(cap) + 1;
  return vecs * (size_t)n * sizeof(double2) + (size_t)cap * (m + 1) * kBlocks * sizeof(double2);
}

GmresSlots::GmresSlots(const DDCheckerboard* dd, int capacity, int restart, double tol, int maxIter)
    : mDD(dd), mCap(capacity), mM(restart), mMaxIter(maxIter), mTol(tol) {
  LSW_REQUIRE(dd != nullptr && dd->deviceReady(), LSW_ERR_STATE, "GmresSlots needs a device DD");
  LSW_REQUIRE(capacity >= 1 && capacity <= kMaxSlots, LSW_ERR_ARG, "slot capacity must be in [1, 64]");
  LSW_REQUIRE(restart >= 1 && restart <= kMaxRestart, LSW_ERR_ARG, "restart length must be in [1, 32]");
  LSW_REQUIRE(maxIter >= 1 && tol >= 0, LSW_ERR_ARG, "bad tolerance or iteration cap");
  mN = dd->N();
  mSlots.resize(mCap);
  const size_t vb = (size_t)mN * sizeof(double2);
  try {
    LSW_HIP(devMalloc(&dV, (size_t)mCap * (mM + 1) * vb));
    LSW_HIP(devMalloc(&dZ, (size_t)mCap * mM * vb));
    LSW_HIP(devMalloc(&dB, (size_t)mCap * vb));
    LSW_HIP(devMalloc(&dX, (size_t)mCap * vb));
    LSW_HIP(devMalloc(&dIn, (size_t)planWidth(mCap) * vb));
    LSW_HIP(devMalloc(&dPZ, (size_t)planWidth(mCap) * vb));
    LSW_HIP(devMalloc(&dW, (size_t)mCap * vb));
    LSW_HIP(devMalloc(&dStage, vb));
    LSW_HIP(devMalloc(&dPart, (size_t)mCap * (mM + 1) * kBlocks * sizeof(double2)));
    LSW_HIP(devMalloc(&dH, (size_t)(2 * mCap * (mM + 1) + mCap) * sizeof(double2)));
  } catch (...) {
    for (double2* p : {dV, dZ, dB, dX, dIn, dPZ, dW, dStage, dPart, dH}) if (p) (void)devFree(p);
    throw;
  }
}

GmresSlots::~GmresSlots() {
  (void)hipStreamSynchronize(mDD->stream());
  IdleFrees idle;  // the engine only runs on its DD's stream
  for (double2* p : {dV, dZ, dB, dX, dIn, dPZ, dW, dStage, dPart, dH}) if (p) (void)devFree(p);
}

int GmresSlots::active() const {
  int k = 0;
  for (const auto& s : mSlots) k += s.state == kActive;
  return k;
}

int GmresSlots::freeSlots() const {
  int k = 0;
  for (const auto& s : mSlots) k += s.state == kFree;
  return k;
}

double GmresSlots::norm(const double2* v) const {
  hipStream_t s = mDD->stream();
  blas::norm2sq(v, mN, dPart, dH, s);
  double2 h;
  LSW_HIP(hipMemcpyAsync(&h, dH, sizeof(double2), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipStreamSynchronize(s));
  return std::sqrt(h.x);
}

void GmresSlots::startCycle(int k, const double2* r, double rnorm) {
  Slot& sl = mSlots[k];
  hipLaunchKernelGGL(k_sscale, dim3(gridN(mN)), dim3(256), 0, mDD->stream(), V(k, 0), r, 1.0 / rnorm, mN);
  sl.j = 0;
  sl.H.assign((size_t)(mM + 1) * mM, 0.0);
  sl.cs.assign(mM, 0.0);
  sl.sn.assign(mM, 0.0);
  sl.g.assign(mM + 1, 0.0);
  sl.g[0] = rnorm;
  sl.state = kActive;
}

int GmresSlots::admit(const double2* b, long long tag) {
  int k = -1;
  for (int i = 0; i < mCap; ++i)
    if (mSlots[i].state == kFree) {
      k = i;
      break;
    }
  LSW_REQUIRE(k >= 0, LSW_ERR_CAPACITY, "GmresSlots: no free slot");
  hipStream_t s = mDD->stream();
  Slot& sl = mSlots[k];
  sl = Slot();
  sl.tag = tag;
  LSW_HIP(hipMemcpyAsync(dB + (long long)k * mN, b, (size_t)mN * sizeof(double2), hipMemcpyDeviceToDevice, s));
  LSW_HIP(hipMemsetAsync(dX + (long long)k * mN, 0, (size_t)mN * sizeof(double2), s));
  sl.bnorm = norm(dB + (long long)k * mN);
  LSW_REQUIRE(std::isfinite(sl.bnorm), LSW_ERR_NUMERIC, "GmresSlots: non-finite right-hand side");
  if (sl.bnorm == 0.0) {
    sl.state = kFinished;  // x = 0 is exact; reported by the next step()
    sl.zeroPending = true;
    return k;
  }
  startCycle(k, dB + (long long)k * mN, sl.bnorm);
  return k;
}

void GmresSlots::release(int k) {
  LSW_REQUIRE(k >= 0 && k < mCap, LSW_ERR_ARG, "GmresSlots: bad slot");
  mSlots[k].state = kFree;
}

void GmresSlots::endCycle(int k, std::vector<Done>& done) {
  Slot& sl = mSlots[k];
  hipStream_t s = mDD->stream();
  const int je = sl.j;
  // back substitution R y = g on the rotated Hessenberg matrix
  Weights w;
  std::memset(&w, 0, sizeof(w));
  std::vector<std::complex<double>> y(je);
  for (int i = je - 1; i >= 0; --i) {
    std::complex<double> acc = sl.g[i];
    for (int c = i + 1; c < je; ++c) acc -= sl.H[(size_t)c * (mM + 1) + i] * y[c];
    y[i] = acc / sl.H[(size_t)i * (mM + 1) + i];
  }
  for (int i = 0; i < je; ++i) w.y[i] = make_double2(y[i].real(), y[i].imag());
  double2* x = dX + (long long)k * mN;
  if (je > 0) hipLaunchKernelGGL(k_scomb, dim3(gridN(mN)), dim3(256), 0, s, x, Z(k, 0), w, je, mN);
  // true residual in FP64 on the global operator
  mDD->applyADev(x, dStage, s);
  hipLaunchKernelGGL(k_ssub, dim3(gridN(mN)), dim3(256), 0, s, dStage, dB + (long long)k * mN, dStage, mN);
  const double rn = norm(dStage);
  const double rel = rn / sl.bnorm;
  LSW_REQUIRE(std::isfinite(rel), LSW_ERR_NUMERIC, "GmresSlots: non-finite residual");
  if (rel <= mTol || sl.iterations >= mMaxIter || rn == 0.0) {
    sl.state = kFinished;
    done.push_back({k, sl.tag, sl.iterations, rel, sl.estimate});
    return;
  }
  startCycle(k, dStage, rn);  // restart from the true residual
}

void GmresSlots::step(std::vector<Done>& done) {
  hipStream_t s = mDD->stream();
  // zero right-hand sides admitted since the last step
  for (int k = 0; k < mCap; ++k) {
    Slot& sl = mSlots[k];
    if (sl.state == kFinished && sl.zeroPending) {
      done.push_back({k, sl.tag, 0, 0.0, 0.0});
      sl.zeroPending = false;
    }
  }
  SlotList L;
  std::memset(&L, 0, sizeof(L));
  for (int k = 0; k < mCap; ++k)
    if (mSlots[k].state == kActive) {
      L.slot[L.n] = k;
      L.nv[L.n] = mSlots[k].j + 1;
      L.maxnv = std::max(L.maxnv, L.nv[L.n]);
      ++L.n;
    }
  if (L.n == 0) return;
  const int na = L.n;
  const unsigned gx = gridN(mN);
  hipLaunchKernelGGL(k_sgather, dim3(gx, na), dim3(256), 0, s, dIn, dV, L, mM, mN);
  // A width whose plans exist already is used for fewer active slots too:
  // building a narrower width's plans costs tens to hundreds of applications'
  // worth of host time on a large board, and the slots that remain at the
  // end of a problem usually finish within a few steps.
  const int built = mDD->builtWidth(na, planWidth(mCap));
  const int nb = built ? built : planWidth(na);
  if (nb > na)
    LSW_HIP(hipMemsetAsync(dIn + (long long)na * mN, 0, (size_t)(nb - na) * mN * sizeof(double2), s));
  static const bool log = std::getenv("LSW_STEP_LOG") != nullptr;
  auto lap = [&](double& t) {  // (LSW_STEP_LOG: phases timed with a wait after each)
    if (!log) return;
    (void)hipStreamSynchronize(s);
    const double n = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    t = n;
  };
  double tg = 0, tp = 0, ta = 0, tgs = 0;
  lap(tg);
  mDD->applyPrecondDev(dIn, dPZ, s, nb);
  lap(tp);
  hipLaunchKernelGGL(k_sscatter, dim3(gx, na), dim3(256), 0, s, dZ, dPZ, L, mM, mN);
  mDD->applyADevBlock(dPZ, dW, na, s);
  lap(ta);
  ++mBlockApplies;
  // two passes of classical Gram-Schmidt against each slot's own basis
  const int np = na * L.maxnv;
  double2* h1 = dH;
  double2* h2 = dH + (long long)mCap * (mM + 1);
  double2* nr = dH + 2LL * mCap * (mM + 1);
  for (int pass = 0; pass < 2; ++pass) {
    double2* h = pass ? h2 : h1;
    hipLaunchKernelGGL(k_sdot_part, dim3(kBlocks, np), dim3(kThreads), 0, s, dV, dW, L, mM, mN, dPart);
    hipLaunchKernelGGL(k_sfinal, dim3(np), dim3(kThreads), 0, s, dPart, h);
    hipLaunchKernelGGL(k_supdate, dim3(gx, na), dim3(256), 0, s, dV, dW, h, L, mM, mN);
  }
  hipLaunchKernelGGL(k_snorm_part, dim3(kBlocks, na), dim3(kThreads), 0, s, dW, mN, dPart);
  hipLaunchKernelGGL(k_sfinal, dim3(na), dim3(kThreads), 0, s, dPart, nr);
  hipLaunchKernelGGL(k_sscale_to, dim3(gx, na), dim3(256), 0, s, dV, dW, nr, L, mM, mN);
  std::vector<double2> c1(np), c2(np), c3(na);
  LSW_HIP(hipMemcpyAsync(c1.data(), h1, np * sizeof(double2), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipMemcpyAsync(c2.data(), h2, np * sizeof(double2), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipMemcpyAsync(c3.data(), nr, na * sizeof(double2), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipStreamSynchronize(s));
  LSW_HIP(hipGetLastError());
  lap(tgs);
  if (log)
    std::fprintf(stderr, "slots N %lld na %d nb %d maxnv %d precond_ms %.2f scatter_A_ms %.2f gs_ms %.2f\n", (long long)mN,
                 na, nb, L.maxnv, 1e3 * (tp - tg), 1e3 * (ta - tp), 1e3 * (tgs - ta));
  ++mSteps;
  for (int a = 0; a < na; ++a) {
    const int k = L.slot[a];
    Slot& sl = mSlots[k];
    const int j = sl.j;
    auto Hc = [&](int i) -> std::complex<double>& { return sl.H[(size_t)j * (mM + 1) + i]; };
    for (int i = 0; i <= j; ++i) {
      const double2 u = c1[a * L.maxnv + i], v = c2[a * L.maxnv + i];
      Hc(i) = std::complex<double>(u.x + v.x, u.y + v.y);
    }
    Hc(j + 1) = std::sqrt(c3[a].x);
    LSW_REQUIRE(std::isfinite(Hc(j + 1).real()), LSW_ERR_NUMERIC, "GMRES breakdown: non-finite norm");
    for (int i = 0; i < j; ++i) {
      const std::complex<double> t = std::conj(sl.cs[i]) * Hc(i) + std::conj(sl.sn[i]) * Hc(i + 1);
      Hc(i + 1) = -sl.sn[i] * Hc(i) + sl.cs[i] * Hc(i + 1);
      Hc(i) = t;
    }
    const std::complex<double> x0 = Hc(j), x1 = Hc(j + 1);
    const double den = std::sqrt(std::norm(x0) + std::norm(x1));
    if (den == 0.0) {
      sl.cs[j] = 1.0;
      sl.sn[j] = 0.0;
    } else {
      sl.cs[j] = x0 / den;
      sl.sn[j] = x1 / den;
    }
    Hc(j) = std::conj(sl.cs[j]) * x0 + std::conj(sl.sn[j]) * x1;
    Hc(j + 1) = 0.0;
    sl.g[j + 1] = -sl.sn[j] * sl.g[j];
    sl.g[j] = std::conj(sl.cs[j]) * sl.g[j];
    sl.j = j + 1;
    sl.iterations++;
    sl.estimate = std::abs(sl.g[j + 1]) / sl.bnorm;
    if (sl.estimate <= mTol || sl.j == mM || sl.iterations >= mMaxIter || den == 0.0) endCycle(k, done);
  }
}

}  // namespace lsw
