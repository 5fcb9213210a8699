// FactorSet: device assembly, multifrontal factorisation and batched ND
// solves (see lsw_factor.h).
#include "lsw_factor.h"

#include <algorithm>
#include <cstring>

#include "lsw_cplx.h"
#include "lsw_stencil.h"

namespace lsw {

// ------------------------------------------------------------------ kernels

// One thread per DOF: cubic-PML stencil (AssemblerHelmholtz2DFD::
// assembleSystemMatrix, src/AssemblerHelmholtz2DFD.cpp:371-457).
__global__ void k_assemble(const CellDev* __restrict__ cells, const double* __restrict__ m,
                           double2* __restrict__ coef, int* __restrict__ bad) {
  const CellDev& cd = cells[blockIdx.y];
  if (cd.mOff < 0) return;
  const int N = cd.nx * cd.ny;
  Geo2 g;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      g.intDom[a][b] = cd.intDom[a][b];
      g.compDom[a][b] = cd.compDom[a][b];
    }
  for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < N; d += gridDim.x * blockDim.x) {
    const int i = d / cd.ny, j = d - i * cd.ny;
    const double x = cd.x0 + (double)(i + 1) * cd.hx;
    const double y = cd.y0 + (double)(j + 1) * cd.hy;
    double cf[10];
    const bool ok = helmholtzStencil(x, y, g, cd.C, cd.omega, m[cd.mOff + d], cd.hx, cd.hy, cf);
    if (!ok) atomicOr(bad, 1);
    for (int dir = 0; dir < 5; ++dir)
      coef[cd.coefOff + (long long)dir * N + d] = make_double2(cf[2 * dir], cf[2 * dir + 1]);
  }
}

// One workgroup per front: assemble F = [F11 F12; F21 F22] (original entries
// plus the children's Schur complements), then Gauss-Jordan with partial
// pivoting restricted to the e fully-summed rows on [F11 F12 I; F21 F22 0]:
//   top rows    -> [I | F11^{-1} F12 | F11^{-1}]
//   bottom rows -> [0 | S            | -F21 F11^{-1}]
// Writes W = F21 F11^{-1}, K = [F11^{-1} | -F11^{-1} F12] and S.
__global__ __launch_bounds__(256) void k_front_factor(
    const CellDev* __restrict__ cells, const ShapeDev* __restrict__ shapes,
    const FrontTask* __restrict__ fts, const double2* __restrict__ coef,
    double2* __restrict__ factor, double2* __restrict__ schur, double2* __restrict__ ws,
    int* __restrict__ info) {
  extern __shared__ double2 colk[];
  __shared__ double redv[256];
  __shared__ int redi[256];
  const FrontTask ft = fts[blockIdx.x];
  const CellDev& cd = cells[ft.cell];
  const ShapeDev& sh = shapes[cd.shape];
  const NdNodeDev nd = sh.nodes[ft.node];
  const int e = nd.e, b = nd.b, R = e + b, Cc = 2 * e + b;
  double2* M = ws + ft.wsOff;
  const int N = cd.nx * cd.ny;
  const int tid = threadIdx.x, nt = blockDim.x;
  const double2 zero = make_double2(0.0, 0.0);

  for (long long idx = tid; idx < (long long)R * Cc; idx += nt) M[idx] = zero;
  __syncthreads();
  for (int k = tid; k < e; k += nt) M[(long long)k * Cc + e + b + k] = make_double2(1.0, 0.0);
  for (int a = tid; a < nd.asmCnt; a += nt) {
    const AsmEntry en = sh.asmList[nd.asmOff + a];
    const double2 v = coef[cd.coefOff + (long long)en.dir * N + en.dof];
    double2* t = &M[(long long)en.fr * Cc + en.fc];
    *t = cadd(*t, v);
  }
  __syncthreads();
  const int cs[2] = {nd.c1, nd.c2};
  for (int ci = 0; ci < 2; ++ci) {
    const int c = cs[ci];
    if (c < 0) continue;
    const NdNodeDev cn = sh.nodes[c];
    const int bc = cn.b;
    const double2* S = schur + cd.schurOff + cn.sOff;
    const int* up = sh.up + cn.upOff;
    for (long long idx = tid; idx < (long long)bc * bc; idx += nt) {
      const int t1 = (int)(idx / bc), t2 = (int)(idx - (long long)t1 * bc);
      double2* t = &M[(long long)up[t1] * Cc + up[t2]];
      *t = cadd(*t, S[idx]);
    }
    __syncthreads();
  }

  for (int k = 0; k < e; ++k) {
    // pivot: largest modulus in column k among rows k..e-1 (lowest index on ties)
    double best = -1.0;
    int bi = e;
    for (int i = k + tid; i < e; i += nt) {
      const double2 v = M[(long long)i * Cc + k];
      const double a = v.x * v.x + v.y * v.y;
      if (a > best) {
        best = a;
        bi = i;
      }
    }
    redv[tid] = best;
    redi[tid] = bi;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
      if (tid < s) {
        const double o = redv[tid + s];
        const int oi = redi[tid + s];
        if (o > redv[tid] || (o == redv[tid] && oi < redi[tid])) {
          redv[tid] = o;
          redi[tid] = oi;
        }
      }
      __syncthreads();
    }
    const int p = redi[0];
    const double pmag = redv[0];
    __syncthreads();
    if (p != k) {
      for (int j = tid; j < Cc; j += nt) {
        const double2 t = M[(long long)k * Cc + j];
        M[(long long)k * Cc + j] = M[(long long)p * Cc + j];
        M[(long long)p * Cc + j] = t;
      }
    }
    __syncthreads();
    double2 piv = M[(long long)k * Cc + k];
    if (!(pmag > 0.0) || !isfinite(pmag)) {
      if (tid == 0) atomicMax(info, 1);
      piv = make_double2(1.0, 0.0);
    }
    const double2 inv = cinv(piv);
    for (int j = k + 1 + tid; j < Cc; j += nt) {
      double2* t = &M[(long long)k * Cc + j];
      *t = cmul(*t, inv);
    }
    for (int i = tid; i < R; i += nt) colk[i] = (i == k) ? zero : M[(long long)i * Cc + k];
    __syncthreads();
    for (int j = k + 1 + tid; j < Cc; j += nt) {
      const double2 rk = M[(long long)k * Cc + j];
      if (rk.x == 0.0 && rk.y == 0.0) continue;
      for (int i = 0; i < R; ++i) {
        if (i == k) continue;
        double2* t = &M[(long long)i * Cc + j];
        *t = cfms(colk[i], rk, *t);
      }
    }
    for (int i = tid; i < R; i += nt) M[(long long)i * Cc + k] = (i == k) ? make_double2(1.0, 0.0) : zero;
    __syncthreads();
  }

  double2* W = factor + cd.factorOff + nd.fwdOff;
  double2* K = factor + cd.factorOff + nd.bwdOff;
  for (long long idx = tid; idx < (long long)b * e; idx += nt) {
    const int k = (int)(idx / b), t = (int)(idx - (long long)k * b);
    W[idx] = cneg(M[(long long)(e + t) * Cc + e + b + k]);
  }
  for (long long idx = tid; idx < (long long)e * R; idx += nt) {
    const int k = (int)(idx / e), r = (int)(idx - (long long)k * e);
    K[idx] = (k < e) ? M[(long long)r * Cc + e + b + k] : cneg(M[(long long)r * Cc + k]);
  }
  if (b > 0) {
    double2* S = schur + cd.schurOff + nd.sOff;
    for (long long idx = tid; idx < (long long)b * b; idx += nt) {
      const int t1 = (int)(idx / b), t2 = (int)(idx - (long long)t1 * b);
      S[idx] = M[(long long)(e + t1) * Cc + e + t2];
    }
  }
}

__global__ void k_to_f32(const double2* __restrict__ a, float2* __restrict__ o, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    o[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// --------------------------------------------------------------- host side
namespace {
template <typename T>
T* devUpload(const std::vector<T>& h, std::vector<void*>& allocs) {
  if (h.empty()) return nullptr;
  T* d = nullptr;
  LSW_HIP(devMalloc(&d, h.size() * sizeof(T)));
  LSW_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  allocs.push_back(d);
  return d;
}
}  // namespace

FactorSet::~FactorSet() {
  for (void* p : mShapeAllocs) (void)devFree(p);
  if (dCells) (void)devFree(dCells);
  if (dShapes) (void)devFree(dShapes);
  if (dCoef) (void)devFree(dCoef);
  if (dM) (void)devFree(dM);
  if (dFactor) (void)devFree(dFactor);
  if (dFactor32) (void)devFree(dFactor32);
}

int FactorSet::addCell(const CellDev& c, const std::vector<double>& m) {
  LSW_REQUIRE((long long)m.size() == (long long)c.nx * c.ny, LSW_ERR_ARG, "FactorSet: m size mismatch");
  mCells.push_back(c);
  mM.push_back(m);
  mCoefHost.emplace_back();
  return (int)mCells.size() - 1;
}

int FactorSet::addCellWithStencil(const CellDev& c, const std::vector<double2>& coef) {
  LSW_REQUIRE((long long)coef.size() == 5LL * c.nx * c.ny, LSW_ERR_ARG, "FactorSet: stencil size mismatch");
  mCells.push_back(c);
  mM.emplace_back();
  mCoefHost.push_back(coef);
  return (int)mCells.size() - 1;
}

long long FactorSet::factorBytes() const {
  return mFactorEntries * (mPrecision == LSW_FP32 ? (long long)sizeof(float2) : (long long)sizeof(double2));
}

void FactorSet::setup(hipStream_t s) {
  LSW_REQUIRE(!mCells.empty(), LSW_ERR_STATE, "FactorSet: no cells");
  // shapes
  for (auto& c : mCells) {
    LSW_REQUIRE(c.nx >= 1 && c.ny >= 1, LSW_ERR_ARG, "FactorSet: empty grid");
    auto key = std::make_pair(c.nx, c.ny);
    auto it = mShapeOf.find(key);
    if (it == mShapeOf.end()) {
      auto t = std::make_unique<NdTree>();
      t->build(c.nx, c.ny);
      LSW_REQUIRE(t->maxHeight < kMaxHeight, LSW_ERR_ARG, "FactorSet: tree too deep");
      mShapeOf[key] = (int)mTrees.size();
      mTrees.push_back(std::move(t));
      it = mShapeOf.find(key);
    }
    c.shape = it->second;
  }
  // offsets
  long long co = 0, mo = 0, fo = 0;
  mMaxDofs = 0;
  mMaxG = 0;
  mMaxHeight = 0;
  mMaxFront = 0;
  for (size_t i = 0; i < mCells.size(); ++i) {
    CellDev& c = mCells[i];
    const NdTree& t = *mTrees[c.shape];
    const long long N = (long long)c.nx * c.ny;
    c.coefOff = co;
    co += 5 * N;
    if (!mM[i].empty()) {
      c.mOff = mo;
      mo += N;
    } else {
      c.mOff = -1;
    }
    c.factorOff = fo;
    fo += t.factorSize;
    mMaxDofs = std::max(mMaxDofs, N);
    mMaxG = std::max(mMaxG, t.gSize);
    mMaxHeight = std::max(mMaxHeight, t.maxHeight);
    mMaxFront = std::max(mMaxFront, t.maxFront);
  }
  mFactorEntries = fo;
  // shape descriptors
  std::vector<ShapeDev> shapes(mTrees.size());
  for (size_t si = 0; si < mTrees.size(); ++si) {
    const NdTree& t = *mTrees[si];
    ShapeDev& d = shapes[si];
    std::memset(&d, 0, sizeof(d));
    d.nodes = devUpload(t.deviceNodes(), mShapeAllocs);
    d.E = devUpload(t.E, mShapeAllocs);
    d.B = devUpload(t.B, mShapeAllocs);
    d.cmap = devUpload(t.cmap, mShapeAllocs);
    d.up = devUpload(t.up, mShapeAllocs);
    d.asmList = devUpload(t.asmList, mShapeAllocs);
    std::vector<SolveItem> fi, bi;
    for (int h = 0; h <= t.maxHeight; ++h) {
      d.fwdStart[h] = (int)fi.size();
      d.fwdCount[h] = (int)t.fwdItems[h].size();
      fi.insert(fi.end(), t.fwdItems[h].begin(), t.fwdItems[h].end());
      d.bwdStart[h] = (int)bi.size();
      d.bwdCount[h] = (int)t.bwdItems[h].size();
      bi.insert(bi.end(), t.bwdItems[h].begin(), t.bwdItems[h].end());
    }
    d.fwdItems = devUpload(fi, mShapeAllocs);
    d.bwdItems = devUpload(bi, mShapeAllocs);
    d.maxHeight = t.maxHeight;
    d.nodeCount = (int)t.nodes.size();
  }
  // flattened solve descriptors over all shapes
  {
    std::vector<SolveNode> flat;
    std::vector<int> pe, pb, pm;
    mShapeNodeBase.assign(mTrees.size(), 0);
    for (size_t si = 0; si < mTrees.size(); ++si) {
      const NdTree& t = *mTrees[si];
      mShapeNodeBase[si] = (int)flat.size();
      const int eb = (int)pe.size(), bb = (int)pb.size(), mb = (int)pm.size();
      pe.insert(pe.end(), t.E.begin(), t.E.end());
      pb.insert(pb.end(), t.B.begin(), t.B.end());
      pm.insert(pm.end(), t.cmap.begin(), t.cmap.end());
      for (const NdNode& nd : t.nodes) {
        SolveNode sn;
        std::memset(&sn, 0, sizeof(sn));
        sn.fwdOff = nd.fwdOff;
        sn.bwdOff = nd.bwdOff;
        sn.e = nd.e;
        sn.b = nd.b;
        sn.E = eb + nd.eOff;
        sn.B = bb + nd.bOff;
        sn.map = mb + nd.mapOff;
        sn.gOff = nd.gOff;
        sn.g1 = nd.c1 >= 0 ? t.nodes[nd.c1].gOff : -1;
        sn.g2 = nd.c2 >= 0 ? t.nodes[nd.c2].gOff : -1;
        flat.push_back(sn);
      }
    }
    dSolveNodes = devUpload(flat, mShapeAllocs);
    dPoolE = devUpload(pe, mShapeAllocs);
    dPoolB = devUpload(pb, mShapeAllocs);
    dPoolMap = devUpload(pm, mShapeAllocs);
  }
  LSW_HIP(devMalloc(&dShapes, shapes.size() * sizeof(ShapeDev)));
  LSW_HIP(hipMemcpy(dShapes, shapes.data(), shapes.size() * sizeof(ShapeDev), hipMemcpyHostToDevice));
  LSW_HIP(devMalloc(&dCells, mCells.size() * sizeof(CellDev)));
  LSW_HIP(hipMemcpy(dCells, mCells.data(), mCells.size() * sizeof(CellDev), hipMemcpyHostToDevice));
  // stencils
  LSW_HIP(devMalloc(&dCoef, co * sizeof(double2)));
  for (size_t i = 0; i < mCells.size(); ++i)
    if (!mCoefHost[i].empty())
      LSW_HIP(hipMemcpy(dCoef + mCells[i].coefOff, mCoefHost[i].data(),
                        mCoefHost[i].size() * sizeof(double2), hipMemcpyHostToDevice));
  if (mo > 0) {
    LSW_HIP(devMalloc(&dM, mo * sizeof(double)));
    for (size_t i = 0; i < mCells.size(); ++i)
      if (!mM[i].empty())
        LSW_HIP(hipMemcpy(dM + mCells[i].mOff, mM[i].data(), mM[i].size() * sizeof(double),
                          hipMemcpyHostToDevice));
    int* dBad = nullptr;
    LSW_HIP(devMalloc(&dBad, sizeof(int)));
    LSW_HIP(hipMemsetAsync(dBad, 0, sizeof(int), s));
    dim3 grid((unsigned)std::min<long long>((mMaxDofs + 255) / 256, 4096), (unsigned)mCells.size());
    hipLaunchKernelGGL(k_assemble, grid, dim3(256), 0, s, dCells, dM, dCoef, dBad);
    LSW_HIP(hipGetLastError());
    int bad = 0;
    LSW_HIP(hipMemcpyAsync(&bad, dBad, sizeof(int), hipMemcpyDeviceToHost, s));
    LSW_HIP(hipStreamSynchronize(s));
    (void)devFree(dBad);
    LSW_REQUIRE(bad == 0, LSW_ERR_ARG, "PML: Evaluation point out of range!");
  }
}

void FactorSet::factorize(int precision, hipStream_t s) {
  LSW_REQUIRE(dCells != nullptr, LSW_ERR_STATE, "FactorSet: setup() not called");
  LSW_REQUIRE(precision == LSW_FP64 || precision == LSW_FP32, LSW_ERR_ARG, "unknown precision");
  if (!dFactor) LSW_HIP(devMalloc(&dFactor, std::max<long long>(mFactorEntries, 1) * sizeof(double2)));
  // Schur workspace for all cells
  std::vector<long long> schurOff(mCells.size());
  long long so = 0;
  for (size_t i = 0; i < mCells.size(); ++i) {
    schurOff[i] = so;
    so += mTrees[mCells[i].shape]->schurSize;
    mCells[i].schurOff = so - mTrees[mCells[i].shape]->schurSize;
  }
  LSW_HIP(hipMemcpy(dCells, mCells.data(), mCells.size() * sizeof(CellDev), hipMemcpyHostToDevice));
  double2* dSchur = nullptr;
  LSW_HIP(devMalloc(&dSchur, std::max<long long>(so, 1) * sizeof(double2)));
  int* dInfo = nullptr;
  LSW_HIP(devMalloc(&dInfo, sizeof(int)));
  LSW_HIP(hipMemsetAsync(dInfo, 0, sizeof(int), s));
  // front batches per height, bounded by a workspace budget
  const long long budget = 512LL << 20;  // complex entries (8 GiB)
  std::vector<std::vector<FrontTask>> batches;
  std::vector<long long> batchWs;
  long long maxWs = 0;
  for (int h = 0; h <= mMaxHeight; ++h) {
    std::vector<FrontTask> cur;
    long long ws = 0;
    for (size_t ci = 0; ci < mCells.size(); ++ci) {
      const NdTree& t = *mTrees[mCells[ci].shape];
      if (h > t.maxHeight) continue;
      for (int p : t.byHeight[h]) {
        const NdNode& nd = t.nodes[p];
        const long long need = (long long)(nd.e + nd.b) * (2 * nd.e + nd.b);
        if (!cur.empty() && ws + need > budget) {
          batches.push_back(cur);
          batchWs.push_back(ws);
          maxWs = std::max(maxWs, ws);
          cur.clear();
          ws = 0;
        }
        cur.push_back({(int)ci, p, ws});
        ws += need;
      }
    }
    if (!cur.empty()) {
      batches.push_back(cur);
      batchWs.push_back(ws);
      maxWs = std::max(maxWs, ws);
    }
    batches.push_back({});  // height separator (keeps order explicit)
    batchWs.push_back(0);
  }
  double2* dWs = nullptr;
  LSW_HIP(devMalloc(&dWs, std::max<long long>(maxWs, 1) * sizeof(double2)));
  FrontTask* dFt = nullptr;
  size_t ftCap = 0;
  for (size_t bi = 0; bi < batches.size(); ++bi) {
    const auto& bt = batches[bi];
    if (bt.empty()) continue;
    if (bt.size() > ftCap) {
      if (dFt) {
        LSW_HIP(hipStreamSynchronize(s));
        (void)devFree(dFt);
      }
      ftCap = bt.size();
      LSW_HIP(devMalloc(&dFt, ftCap * sizeof(FrontTask)));
    }
    LSW_HIP(hipMemcpyAsync(dFt, bt.data(), bt.size() * sizeof(FrontTask), hipMemcpyHostToDevice, s));
    const size_t shm = (size_t)mMaxFront * sizeof(double2);
    hipLaunchKernelGGL(k_front_factor, dim3((unsigned)bt.size()), dim3(256), shm, s, dCells, dShapes,
                       dFt, dCoef, dFactor, dSchur, dWs, dInfo);
    LSW_HIP(hipGetLastError());
    LSW_HIP(hipStreamSynchronize(s));  // the host task buffer is reused
  }
  LSW_HIP(hipMemcpyAsync(&mInfo, dInfo, sizeof(int), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipStreamSynchronize(s));
  if (dFt) (void)devFree(dFt);
  (void)devFree(dWs);
  (void)devFree(dSchur);
  (void)devFree(dInfo);
  LSW_REQUIRE(mInfo == 0, LSW_ERR_NUMERIC, "factorization failed: zero or non-finite pivot");
  mPrecision = precision;
  if (precision == LSW_FP32) {
    if (!dFactor32) LSW_HIP(devMalloc(&dFactor32, std::max<long long>(mFactorEntries, 1) * sizeof(float2)));
    hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, s, dFactor, dFactor32, mFactorEntries);
    LSW_HIP(hipGetLastError());
    LSW_HIP(hipStreamSynchronize(s));
    (void)devFree(dFactor);
    dFactor = nullptr;
  }
  mFactorized = true;
}

}  // namespace lsw

namespace lsw {
// Assemble the stencil of one grid (used for the global operator).
double2* assembleStencilDev(const CellDev& c, const std::vector<double>& m, hipStream_t s) {
  const long long N = (long long)c.nx * c.ny;
  LSW_REQUIRE((long long)m.size() == N, LSW_ERR_ARG, "assembleStencilDev: m size mismatch");
  CellDev cd = c;
  cd.mOff = 0;
  cd.coefOff = 0;
  CellDev* dC = nullptr;
  double* dM = nullptr;
  double2* dCoef = nullptr;
  int* dBad = nullptr;
  LSW_HIP(devMalloc(&dC, sizeof(CellDev)));
  LSW_HIP(devMalloc(&dM, N * sizeof(double)));
  LSW_HIP(devMalloc(&dCoef, 5 * N * sizeof(double2)));
  LSW_HIP(devMalloc(&dBad, sizeof(int)));
  LSW_HIP(hipMemcpy(dC, &cd, sizeof(CellDev), hipMemcpyHostToDevice));
  LSW_HIP(hipMemcpy(dM, m.data(), N * sizeof(double), hipMemcpyHostToDevice));
  LSW_HIP(hipMemsetAsync(dBad, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_assemble, dim3((unsigned)std::min<long long>((N + 255) / 256, 8192), 1), dim3(256), 0, s,
                     dC, dM, dCoef, dBad);
  LSW_HIP(hipGetLastError());
  int bad = 0;
  LSW_HIP(hipMemcpyAsync(&bad, dBad, sizeof(int), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipStreamSynchronize(s));
  (void)devFree(dC);
  (void)devFree(dM);
  (void)devFree(dBad);
  if (bad) {
    (void)devFree(dCoef);
    throw Error(LSW_ERR_ARG, "PML: Evaluation point out of range!");
  }
  return dCoef;
}
}  // namespace lsw
