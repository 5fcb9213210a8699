// Batched nested-dissection solves: the hot loop of the L-sweeps.
//
// One solve of a cell is, per tree height, a set of dense matrix-vector
// products (forward: g_p = W_p r_E + children updates; backward:
// x_E = [F11^{-1} | -F11^{-1} F12] [r_E; x_B]).  A launch covers one height of
// every task of a sweep step.  Work lists are flat {task, node, row chunk}
// entries prepared on the host once per step:
//   * the forward list only holds nodes whose subtree region meets the
//     right-hand side support (sweep right-hand sides live on 2-4 grid lines),
//   * the backward list holds every node, and skips the F11^{-1} r_E columns of
//     nodes the forward pass did not visit (their r_E is zero),
//   * a workgroup walks `ipb` consecutive entries so that tiny leaf fronts do
//     not pay a workgroup launch each;
//   * heights with few, large entries (the separators near the root, where a
//     front step holds only a few dozen of them) give each entry W = 2..8
//     wavefronts that split its columns, so the launch still keeps enough
//     loads in flight to approach the HBM bandwidth.
// Factors are column-major, so the lanes of a wavefront read consecutive rows
// of one column (coalesced); right-hand-side values are broadcast from LDS.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "lsw_cplx.h"
#include "lsw_factor.h"

namespace lsw {

namespace {
__device__ __forceinline__ bool bitSet(const unsigned* __restrict__ a, int node) {
  return (a[node >> 5] >> (node & 31)) & 1u;
}

// y[row] = sum_k M[k*ld + row] * v[k] over k in [k0, K) for rows [0, rows).
// A workgroup of W wavefronts works on one product: a wavefront splits into
// RP rows x CG column groups, the W*CG column groups stride over the columns
// (8 independent loads in flight per lane), partial sums are reduced across
// the column groups of a wavefront by lane shuffles and across wavefronts in
// a fixed order through LDS (`red`: W*64*NR entries).  On return the owner
// lanes (wavefront 0, first column group, r < rows) hold the result.
template <typename T2, int NR, int W>
__device__ __forceinline__ bool gemvCols(const T2* __restrict__ M, long long ld, int rows, int k0, int K,
                                         const T2* v, int vstride, T2* red, T2 (&acc)[NR], int& r) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int RP = 1;
  while (RP < rows) RP <<= 1;
  const int CG = 64 / RP;
  r = lane & (RP - 1);
  const int cg = lane / RP;
  const int CGA = W * CG;
#pragma unroll
  for (int q = 0; q < NR; ++q) acc[q] = czero<T2>();
  if (r < rows) {
    const T2* p = M + r;
    int k = k0 + w * CG + cg;
    for (; k + 7 * CGA < K; k += 8 * CGA) {
      T2 m[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) m[u] = p[(long long)(k + u * CGA) * ld];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = cfma(m[u], v[q * vstride + k + u * CGA], acc[q]);
    }
    for (; k < K; k += CGA) {
      const T2 m0 = p[(long long)k * ld];
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = cfma(m0, v[q * vstride + k], acc[q]);
    }
  }
  for (int off = RP; off < 64; off <<= 1) {
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = cadd(acc[q], shfl_xor2(acc[q], off));
  }
  if constexpr (W == 1) {
    return cg == 0 && r < rows;
  } else {
    if (cg == 0 && r < rows) {
#pragma unroll
      for (int q = 0; q < NR; ++q) red[(w * 64 + r) * NR + q] = acc[q];
    }
    __syncthreads();
    const bool own = w == 0 && cg == 0 && r < rows;
    if (own) {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        T2 s = red[r * NR + q];
        for (int ww = 1; ww < W; ++ww) s = cadd(s, red[(ww * 64 + r) * NR + q]);
        acc[q] = s;
      }
    }
    return own;
  }
}
}  // namespace

// LDS layout of both kernels: v (NR x cols right-hand-side values), then the
// cross-wavefront reduction buffer (W x 64 x NR) when W > 1.
template <typename T2, int NR, int W>
__global__ __launch_bounds__(64 * W) void k_nd_fwd(const CellDev* __restrict__ cells,
                                                   const ShapeDev* __restrict__ shapes,
                                                   const SolveTask* __restrict__ tasks,
                                                   const int4* __restrict__ ent, int nent, int ipb, int cols,
                                                   const unsigned* __restrict__ active, int aw,
                                                   const T2* __restrict__ factor, T2* __restrict__ X,
                                                   T2* __restrict__ Rb, T2* __restrict__ G, long long xs,
                                                   long long gs) {
  extern __shared__ char smem_raw[];
  T2* v = reinterpret_cast<T2*>(smem_raw);
  T2* red = v + (size_t)NR * cols;
  const int tid = threadIdx.x;
  for (int q = 0; q < ipb; ++q) {
    const int idx = blockIdx.x * ipb + q;
    if (idx >= nent) break;
    const int4 en = ent[idx];
    const SolveTask& tk = tasks[en.x];
    const CellDev& cd = cells[tk.cell];
    const ShapeDev& sh = shapes[cd.shape];
    const NdNodeDev nd = sh.nodes[en.y];
    const int e = nd.e, b = nd.b, F = e + b;
    const unsigned* act = active + (long long)en.x * aw;
    const int* E = sh.E + nd.eOff;
    const int* m1 = sh.cmap + nd.mapOff;
    const int* m2 = m1 + F;
    const bool a1ok = nd.c1 >= 0 && bitSet(act, nd.c1);
    const bool a2ok = nd.c2 >= 0 && bitSet(act, nd.c2);
    const int g1 = a1ok ? sh.nodes[nd.c1].gOff : 0;
    const int g2 = a2ok ? sh.nodes[nd.c2].gOff : 0;
    const int nr = tk.nrhs;
    for (int k = tid; k < e; k += 64 * W) {
      const int a1 = a1ok ? m1[k] : -1, a2 = a2ok ? m2[k] : -1, dof = E[k];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r >= nr) break;
        const long long sl = tk.slot[r];
        T2 val = X[sl * xs + dof];
        if (a1 >= 0) val = csub(val, G[sl * gs + g1 + a1]);
        if (a2 >= 0) val = csub(val, G[sl * gs + g2 + a2]);
        v[r * e + k] = val;
        if (en.z == 0) Rb[sl * xs + dof] = val;
      }
    }
    __syncthreads();
    const int rows = en.w;
    if (rows > 0) {
      T2 acc[NR];
      int r;
      const bool own = gemvCols<T2, NR, W>(factor + cd.factorOff + nd.fwdOff + en.z, b, rows, 0, e, v, e, red,
                                           acc, r);
      if (own) {
        const int f = e + en.z + r;
        const int a1 = a1ok ? m1[f] : -1, a2 = a2ok ? m2[f] : -1;
#pragma unroll
        for (int qq = 0; qq < NR; ++qq) {
          if (qq >= nr) break;
          const long long sl = tk.slot[qq];
          T2 g = acc[qq];
          if (a1 >= 0) g = cadd(g, G[sl * gs + g1 + a1]);
          if (a2 >= 0) g = cadd(g, G[sl * gs + g2 + a2]);
          G[sl * gs + nd.gOff + en.z + r] = g;
        }
      }
    }
    __syncthreads();
  }
}

template <typename T2, int NR, int W>
__global__ __launch_bounds__(64 * W) void k_nd_bwd(const CellDev* __restrict__ cells,
                                                   const ShapeDev* __restrict__ shapes,
                                                   const SolveTask* __restrict__ tasks,
                                                   const int4* __restrict__ ent, int nent, int ipb, int cols,
                                                   const unsigned* __restrict__ active, int aw,
                                                   const T2* __restrict__ factor, T2* __restrict__ X,
                                                   const T2* __restrict__ Rb, long long xs) {
  extern __shared__ char smem_raw[];
  T2* v = reinterpret_cast<T2*>(smem_raw);
  T2* red = v + (size_t)NR * cols;
  const int tid = threadIdx.x;
  for (int q = 0; q < ipb; ++q) {
    const int idx = blockIdx.x * ipb + q;
    if (idx >= nent) break;
    const int4 en = ent[idx];
    const SolveTask& tk = tasks[en.x];
    const CellDev& cd = cells[tk.cell];
    const ShapeDev& sh = shapes[cd.shape];
    const NdNodeDev nd = sh.nodes[en.y];
    const int e = nd.e, b = nd.b, F = e + b;
    const bool visited = bitSet(active + (long long)en.x * aw, en.y);
    const int k0 = visited ? 0 : e;
    const int* E = sh.E + nd.eOff;
    const int* Bp = sh.B + nd.bOff;
    const int nr = tk.nrhs;
    for (int k = k0 + tid; k < F; k += 64 * W) {
#pragma unroll
      for (int qq = 0; qq < NR; ++qq) {
        if (qq >= nr) break;
        const long long sl = tk.slot[qq];
        v[qq * F + k] = (k < e) ? Rb[sl * xs + E[k]] : X[sl * xs + Bp[k - e]];
      }
    }
    __syncthreads();
    T2 acc[NR];
    int r;
    const int rows = en.w;
    const bool own = gemvCols<T2, NR, W>(factor + cd.factorOff + nd.bwdOff + en.z, e, rows, k0, F, v, F, red,
                                         acc, r);
    if (own) {
      const int dof = E[en.z + r];
#pragma unroll
      for (int qq = 0; qq < NR; ++qq) {
        if (qq >= nr) break;
        X[(long long)tk.slot[qq] * xs + dof] = acc[qq];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ plans
SolvePlan::~SolvePlan() {
  if (dTasks) (void)hipFree(dTasks);
  if (dEntries) (void)hipFree(dEntries);
  if (dActive) (void)hipFree(dActive);
}

namespace {
constexpr double kMaxLds = 65536.0;  // bytes of dynamic LDS per workgroup

bool meets(const NdNode& nd, const std::vector<Rect>& sup) {
  if (sup.empty()) return true;
  for (const Rect& r : sup)
    if (r.i0 < nd.i1 && nd.i0 < r.i1 && r.j0 < nd.j1 && nd.j0 < r.j1) return true;
  return false;
}
}  // namespace

void FactorSet::buildPlan(const std::vector<SolveTask>& tasks, const std::vector<std::vector<Rect>>& supports,
                          SolvePlan& plan) const {
  LSW_REQUIRE(!mTrees.empty(), LSW_ERR_STATE, "FactorSet: setup() not called");
  LSW_REQUIRE(supports.empty() || supports.size() == tasks.size(), LSW_ERR_ARG, "one support per task");
  const int nt = (int)tasks.size();
  LSW_REQUIRE(nt > 0, LSW_ERR_ARG, "empty plan");
  int maxNodes = 0;
  for (auto& t : mTrees) maxNodes = std::max(maxNodes, (int)t->nodes.size());
  plan.activeWords = (maxNodes + 31) / 32;
  plan.ntasks = nt;
  plan.hTasks = tasks;
  plan.maxNrhs = 1;
  std::vector<unsigned> act((size_t)nt * plan.activeWords, 0u);
  const double S = mPrecision == LSW_FP32 ? 8.0 : 16.0;
  // entries per height
  std::vector<std::vector<int4>> fw(mMaxHeight + 1), bw(mMaxHeight + 1);
  std::vector<int> fcols(mMaxHeight + 1, 1), bcols(mMaxHeight + 1, 1);
  std::vector<double> fbytes(mMaxHeight + 1, 0.0), bbytes(mMaxHeight + 1, 0.0);
  std::vector<double> fwork(mMaxHeight + 1, 0.0), bwork(mMaxHeight + 1, 0.0);
  for (int t = 0; t < nt; ++t) {
    const SolveTask& tk = tasks[t];
    LSW_REQUIRE(tk.nrhs >= 1 && tk.nrhs <= kMaxRhs, LSW_ERR_ARG, "bad task right-hand-side count");
    LSW_REQUIRE(tk.cell >= 0 && tk.cell < (int)mCells.size(), LSW_ERR_ARG, "bad task cell");
    plan.maxNrhs = std::max(plan.maxNrhs, tk.nrhs);
    const NdTree& tr = *mTrees[mCells[tk.cell].shape];
    const std::vector<Rect> none;
    const std::vector<Rect>& sup = supports.empty() ? none : supports[t];
    const double r = tk.nrhs;
    for (int p = 0; p < (int)tr.nodes.size(); ++p) {
      const NdNode& nd = tr.nodes[p];
      const bool on = meets(nd, sup);
      if (on) act[(size_t)t * plan.activeWords + (p >> 5)] |= 1u << (p & 31);
      const int h = nd.height;
      const double e = nd.e, b = nd.b;
      if (on) {
        if (nd.b == 0) {
          fw[h].push_back(make_int4(t, p, 0, 0));
        } else {
          for (int r0 = 0; r0 < nd.b; r0 += kSolveRows)
            fw[h].push_back(make_int4(t, p, r0, std::min(kSolveRows, nd.b - r0)));
        }
        fcols[h] = std::max(fcols[h], nd.e);
        fbytes[h] += S * (b * e + r * (2 * e + 2 * b));
        fwork[h] += b * e + e;
      }
      for (int r0 = 0; r0 < nd.e; r0 += kSolveRows)
        bw[h].push_back(make_int4(t, p, r0, std::min(kSolveRows, nd.e - r0)));
      bcols[h] = std::max(bcols[h], nd.e + nd.b);
      const double kcols = on ? e + b : b;
      bbytes[h] += S * (e * kcols + r * (kcols + e));
      bwork[h] += e * kcols;
    }
  }
  // flatten: forward heights ascending, backward descending
  std::vector<int4> all;
  auto addLaunch = [&](std::vector<SolvePlan::Launch>& out, const std::vector<int4>& v, int cols, double bytes,
                       double work) {
    if (v.empty()) return;
    SolvePlan::Launch L;
    L.off = (long long)all.size();
    L.n = (int)v.size();
    L.cols = cols;
    L.bytes = bytes;
    // several entries per workgroup when entries are small, while keeping
    // enough workgroups to fill the device
    const double per = work / v.size();
    int ipb = (int)std::max(1.0, std::min(16.0, 2048.0 / std::max(per, 1.0)));
    while (ipb > 1 && (v.size() + ipb - 1) / ipb < 4096) ipb /= 2;
    L.ipb = ipb;
    // few, large entries: several wavefronts per entry split the columns
    int W = 1;
    while (W < 8 && (double)v.size() * W * 2 <= 4096.0 && per >= 1024.0 * W * 2) W *= 2;
    // LDS: NR x cols right-hand-side values + W x 64 x NR reduction entries
    const double nrk = plan.maxNrhs == 1 ? 1.0 : (double)kMaxRhs;
    auto lds = [&](int w) { return S * nrk * ((double)cols + (w > 1 ? 64.0 * w : 0.0)); };
    while (W > 1 && lds(W) > kMaxLds) W /= 2;
    LSW_REQUIRE(lds(1) <= kMaxLds, LSW_ERR_CAPACITY, "front too wide for the solve kernel's LDS staging");
    L.waves = W;
    if (W > 1) L.ipb = 1;
    all.insert(all.end(), v.begin(), v.end());
    out.push_back(L);
  };
  plan.fwd.clear();
  plan.bwd.clear();
  plan.fwdBytes = plan.bwdBytes = 0;
  for (int h = 0; h <= mMaxHeight; ++h) {
    addLaunch(plan.fwd, fw[h], fcols[h], fbytes[h], fwork[h]);
    plan.fwdBytes += fbytes[h];
  }
  for (int h = mMaxHeight; h >= 0; --h) {
    addLaunch(plan.bwd, bw[h], bcols[h], bbytes[h], bwork[h]);
    plan.bwdBytes += bbytes[h];
  }
  if (std::getenv("LSW_PLAN_LOG")) {
    for (int pass = 0; pass < 2; ++pass)
      for (const auto& L : pass ? plan.bwd : plan.fwd)
        std::fprintf(stderr, "plan %s tasks %d nrhs %d: entries %d ipb %d waves %d cols %d bytes %.0f\n",
                     pass ? "bwd" : "fwd", nt, plan.maxNrhs, L.n, L.ipb, L.waves, L.cols, L.bytes);
  }
  if (plan.dTasks) (void)hipFree(plan.dTasks);
  if (plan.dEntries) (void)hipFree(plan.dEntries);
  if (plan.dActive) (void)hipFree(plan.dActive);
  LSW_HIP(hipMalloc(&plan.dTasks, tasks.size() * sizeof(SolveTask)));
  LSW_HIP(hipMemcpy(plan.dTasks, tasks.data(), tasks.size() * sizeof(SolveTask), hipMemcpyHostToDevice));
  LSW_HIP(hipMalloc(&plan.dEntries, std::max<size_t>(1, all.size()) * sizeof(int4)));
  LSW_HIP(hipMemcpy(plan.dEntries, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice));
  LSW_HIP(hipMalloc(&plan.dActive, act.size() * sizeof(unsigned)));
  LSW_HIP(hipMemcpy(plan.dActive, act.data(), act.size() * sizeof(unsigned), hipMemcpyHostToDevice));
}

namespace {
struct LaunchTimer {
  SolveProfile* prof;
  hipStream_t s;
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
  std::vector<double> bytes;
  void begin() {
    if (!prof) return;
    hipEvent_t e;
    LSW_HIP(hipEventCreate(&e));
    LSW_HIP(hipEventRecord(e, s));
    ev.push_back(e);
  }
  void end(int k, double b) {
    if (!prof) return;
    hipEvent_t e;
    LSW_HIP(hipEventCreate(&e));
    LSW_HIP(hipEventRecord(e, s));
    ev.push_back(e);
    kind.push_back(k);
    bytes.push_back(b);
  }
  void finish() {
    if (!prof) return;
    LSW_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < kind.size(); ++i) {
      float ms = 0;
      LSW_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      if (kind[i] == 0) {
        prof->fwdMs += ms;
        prof->fwdLaunches++;
        prof->fwdBytes += bytes[i];
      } else {
        prof->bwdMs += ms;
        prof->bwdLaunches++;
        prof->bwdBytes += bytes[i];
      }
    }
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

template <typename T2, int NR, int W>
void launchFwd(const SolvePlan& plan, const SolvePlan::Launch& L, const CellDev* dCells, const ShapeDev* dShapes,
               const T2* factor, T2* X, T2* R, T2* G, long long xs, long long gs, hipStream_t s) {
  const unsigned grid = (unsigned)((L.n + L.ipb - 1) / L.ipb);
  const size_t shm = ((size_t)NR * L.cols + (W > 1 ? (size_t)W * 64 * NR : 0)) * sizeof(T2);
  hipLaunchKernelGGL((k_nd_fwd<T2, NR, W>), dim3(grid), dim3(64 * W), shm, s, dCells, dShapes, plan.dTasks,
                     plan.dEntries + L.off, L.n, L.ipb, L.cols, plan.dActive, plan.activeWords, factor, X, R, G, xs,
                     gs);
}

template <typename T2, int NR, int W>
void launchBwd(const SolvePlan& plan, const SolvePlan::Launch& L, const CellDev* dCells, const ShapeDev* dShapes,
               const T2* factor, T2* X, T2* R, long long xs, hipStream_t s) {
  const unsigned grid = (unsigned)((L.n + L.ipb - 1) / L.ipb);
  const size_t shm = ((size_t)NR * L.cols + (W > 1 ? (size_t)W * 64 * NR : 0)) * sizeof(T2);
  hipLaunchKernelGGL((k_nd_bwd<T2, NR, W>), dim3(grid), dim3(64 * W), shm, s, dCells, dShapes, plan.dTasks,
                     plan.dEntries + L.off, L.n, L.ipb, L.cols, plan.dActive, plan.activeWords, factor, X, R, xs);
}

template <typename T2, int NR>
void runPlan(const SolvePlan& plan, const CellDev* dCells, const ShapeDev* dShapes, const T2* factor, T2* X, T2* R,
             T2* G, long long xs, long long gs, hipStream_t s, SolveProfile* prof) {
  LaunchTimer tm{prof, s, {}, {}, {}};
  for (const auto& L : plan.fwd) {
    tm.begin();
    switch (L.waves) {
      case 1: launchFwd<T2, NR, 1>(plan, L, dCells, dShapes, factor, X, R, G, xs, gs, s); break;
      case 2: launchFwd<T2, NR, 2>(plan, L, dCells, dShapes, factor, X, R, G, xs, gs, s); break;
      case 4: launchFwd<T2, NR, 4>(plan, L, dCells, dShapes, factor, X, R, G, xs, gs, s); break;
      default: launchFwd<T2, NR, 8>(plan, L, dCells, dShapes, factor, X, R, G, xs, gs, s); break;
    }
    tm.end(0, L.bytes);
  }
  for (const auto& L : plan.bwd) {
    tm.begin();
    switch (L.waves) {
      case 1: launchBwd<T2, NR, 1>(plan, L, dCells, dShapes, factor, X, R, xs, s); break;
      case 2: launchBwd<T2, NR, 2>(plan, L, dCells, dShapes, factor, X, R, xs, s); break;
      case 4: launchBwd<T2, NR, 4>(plan, L, dCells, dShapes, factor, X, R, xs, s); break;
      default: launchBwd<T2, NR, 8>(plan, L, dCells, dShapes, factor, X, R, xs, s); break;
    }
    tm.end(1, L.bytes);
  }
  LSW_HIP(hipGetLastError());
  tm.finish();
}
}  // namespace

void FactorSet::solve(const SolvePlan& plan, void* X, void* R, void* G, long long xs, long long gs,
                      hipStream_t s, SolveProfile* prof) const {
  LSW_REQUIRE(mFactorized, LSW_ERR_STATE, "solve called before factorize");
  if (plan.ntasks <= 0) return;
  if (mPrecision == LSW_FP64) {
    if (plan.maxNrhs == 1)
      runPlan<double2, 1>(plan, dCells, dShapes, dFactor, (double2*)X, (double2*)R, (double2*)G, xs, gs, s, prof);
    else
      runPlan<double2, kMaxRhs>(plan, dCells, dShapes, dFactor, (double2*)X, (double2*)R, (double2*)G, xs, gs, s,
                                prof);
  } else {
    if (plan.maxNrhs == 1)
      runPlan<float2, 1>(plan, dCells, dShapes, dFactor32, (float2*)X, (float2*)R, (float2*)G, xs, gs, s, prof);
    else
      runPlan<float2, kMaxRhs>(plan, dCells, dShapes, dFactor32, (float2*)X, (float2*)R, (float2*)G, xs, gs, s,
                               prof);
  }
}

}  // namespace lsw
