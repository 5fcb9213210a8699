// FactorSet: device assembly, multifrontal factorisation and batched ND
// solves (see lsw_factor.h).
#include "lsw_factor.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "lsw_cplx.h"
#include "lsw_stencil.h"

namespace lsw {

// ------------------------------------------------------------------ kernels

// One thread per DOF: cubic-PML stencil (AssemblerHelmholtz2DFD::
// assembleSystemMatrix, src/AssemblerHelmholtz2DFD.cpp:371-457).
__global__ void k_assemble(const CellDev* __restrict__ cells, const double* __restrict__ m,
                           double2* __restrict__ coef) {
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
    // (every DOF inside the PML's domain: checked on the host, gridInsidePml)
    (void)helmholtzStencil(x, y, g, cd.C, cd.omega, m[cd.mOff + d], cd.hx, cd.hy, cf);
    for (int dir = 0; dir < 5; ++dir)
      coef[cd.coefOff + (long long)dir * N + d] = make_double2(cf[2 * dir], cf[2 * dir + 1]);
  }
}

// The dense front of node p, in the workspace, row-major R x Cc with
// R = e + b and Cc = 2e + b: [F11 F12 I; F21 F22 0] (original entries plus
// the children's Schur complements).  Gauss-Jordan with partial pivoting
// restricted to the e fully-summed rows turns it into
//   top rows    -> [I | F11^{-1} F12 | F11^{-1}]
//   bottom rows -> [0 | S            | -F21 F11^{-1}]
// from which W = F21 F11^{-1}, K = [F11^{-1} | -F11^{-1} F12] and S are
// written.  Two schedules: one workgroup per front doing all of it (kept
// for tuning), and the blocked one used by default: panels of kPanel pivots
// (one workgroup per front) each followed by a batched rank-kPanel update of
// the rest of every front of the height, tiled over many workgroups (the
// bulk of the arithmetic, and no longer one CU per front).
template <typename W2>
struct FrontView {
  NdNodeDev nd;
  int e, b, R, Cc, N;
  W2* M;
};

template <typename W2>
__device__ __forceinline__ FrontView<W2> frontView(const CellDev* __restrict__ cells,
                                                   const ShapeDev* __restrict__ shapes, const FrontTask& ft,
                                                   W2* __restrict__ ws) {
  FrontView<W2> f;
  const CellDev& cd = cells[ft.cell];
  const ShapeDev& sh = shapes[cd.shape];
  f.nd = sh.nodes[ft.node];
  f.e = f.nd.e;
  f.b = f.nd.b;
  f.R = f.e + f.b;
  f.Cc = 2 * f.e + f.b;
  f.N = cd.nx * cd.ny;
  f.M = ws + ft.wsOff;
  return f;
}

template <typename W2>
__device__ __forceinline__ W2 cinvT(W2 a) {
  using R = typename CTraits<W2>::R;
  const R d = a.x * a.x + a.y * a.y;
  return CTraits<W2>::make(a.x / d, -a.y / d);
}

// W2: the arithmetic; O2: the factor and Schur-complement storage (a
// mixed-precision factorisation computes some heights in FP32 and keeps FP64
// storage for the heights above).
template <typename O2, typename W2>
__device__ __forceinline__ O2 ccast(W2 v) {
  return CTraits<O2>::make((typename CTraits<O2>::R)v.x, (typename CTraits<O2>::R)v.y);
}

// zero the front, put the identity block, add the original entries (the
// FP64 stencil, rounded to the work type) and the children's Schur
// complements (stored as O2; one workgroup)
template <typename W2, typename O2>
__device__ void frontAssemble(const CellDev& cd, const ShapeDev& sh, const FrontView<W2>& f,
                              const double2* __restrict__ coef, const O2* __restrict__ schur) {
  using R = typename CTraits<W2>::R;
  const int tid = threadIdx.x, nt = blockDim.x;
  const W2 zero = czero<W2>();
  W2* M = f.M;
  const int e = f.e, Cc = f.Cc, b = f.b;
  for (long long idx = tid; idx < (long long)f.R * Cc; idx += nt) M[idx] = zero;
  __syncthreads();
  for (int k = tid; k < e; k += nt) M[(long long)k * Cc + e + b + k] = CTraits<W2>::make(1, 0);
  for (int a = tid; a < f.nd.asmCnt; a += nt) {
    const AsmEntry en = sh.asmList[f.nd.asmOff + a];
    const double2 v = coef[cd.coefOff + (long long)en.dir * f.N + en.dof];
    W2* t = &M[(long long)en.fr * Cc + en.fc];
    *t = cadd(*t, CTraits<W2>::make((R)v.x, (R)v.y));
  }
  __syncthreads();
  const int cs[2] = {f.nd.c1, f.nd.c2};
  for (int ci = 0; ci < 2; ++ci) {
    const int c = cs[ci];
    if (c < 0) continue;
    const NdNodeDev cn = sh.nodes[c];
    const int bc = cn.b;
    const O2* S = schur + cd.schurOff + cn.sOff;
    const int* up = sh.up + cn.upOff;
    for (long long idx = tid; idx < (long long)bc * bc; idx += nt) {
      const int t1 = (int)(idx / bc), t2 = (int)(idx - (long long)t1 * bc);
      W2* t = &M[(long long)up[t1] * Cc + up[t2]];
      *t = cadd(*t, ccast<W2>(S[idx]));
    }
    __syncthreads();
  }
}

// write W, K and S from the reduced front (one workgroup)
template <typename W2, typename O2>
__device__ void frontStore(const CellDev& cd, const FrontView<W2>& f, O2* __restrict__ factor,
                           O2* __restrict__ schur) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const W2* M = f.M;
  const int e = f.e, b = f.b, R = f.R, Cc = f.Cc;
  O2* W = factor + cd.factorOff + f.nd.fwdOff;
  O2* K = factor + cd.factorOff + f.nd.bwdOff;
  for (long long idx = tid; idx < (long long)b * e; idx += nt) {
    const int k = (int)(idx / b), t = (int)(idx - (long long)k * b);
    W[idx] = ccast<O2>(cneg(M[(long long)(e + t) * Cc + e + b + k]));
  }
  for (long long idx = tid; idx < (long long)e * R; idx += nt) {
    const int k = (int)(idx / e), r = (int)(idx - (long long)k * e);
    K[idx] = ccast<O2>((k < e) ? M[(long long)r * Cc + e + b + k] : cneg(M[(long long)r * Cc + k]));
  }
  if (b > 0) {
    O2* S = schur + cd.schurOff + f.nd.sOff;
    for (long long idx = tid; idx < (long long)b * b; idx += nt) {
      const int t1 = (int)(idx / b), t2 = (int)(idx - (long long)t1 * b);
      S[idx] = ccast<O2>(M[(long long)(e + t1) * Cc + e + t2]);
    }
  }
}

// Pivot of column k: the largest modulus among rows k..e-1 (lowest index on
// ties), reduced over the workgroup (256 threads).
template <typename W2>
__device__ __forceinline__ int pivotRow(const W2* __restrict__ M, int Cc, int k, int e, double* redv, int* redi,
                                        double& pmag) {
  const int tid = threadIdx.x, nt = blockDim.x;
  double best = -1.0;
  int bi = e;
  for (int i = k + tid; i < e; i += nt) {
    const W2 v = M[(long long)i * Cc + k];
    const double a = (double)(v.x * v.x + v.y * v.y);
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
  pmag = redv[0];
  __syncthreads();
  return p;
}

// One workgroup per front: assemble, full Gauss-Jordan, store.
template <typename W2>
__global__ __launch_bounds__(256) void k_front_factor(
    const CellDev* __restrict__ cells, const ShapeDev* __restrict__ shapes,
    const FrontTask* __restrict__ fts, const double2* __restrict__ coef,
    W2* __restrict__ factor, W2* __restrict__ schur, W2* __restrict__ ws,
    int* __restrict__ info) {
  extern __shared__ double2 lswDynLds[];
  W2* colk = reinterpret_cast<W2*>(lswDynLds);
  __shared__ double redv[256];
  __shared__ int redi[256];
  const FrontTask ft = fts[blockIdx.x];
  const CellDev& cd = cells[ft.cell];
  const ShapeDev& sh = shapes[cd.shape];
  const FrontView<W2> f = frontView(cells, shapes, ft, ws);
  frontAssemble(cd, sh, f, coef, schur);
  W2* M = f.M;
  const int e = f.e, R = f.R, Cc = f.Cc;
  const int tid = threadIdx.x, nt = blockDim.x;
  const W2 zero = czero<W2>();
  for (int k = 0; k < e; ++k) {
    double pmag;
    const int p = pivotRow(M, Cc, k, e, redv, redi, pmag);
    if (p != k) {
      for (int j = tid; j < Cc; j += nt) {
        const W2 t = M[(long long)k * Cc + j];
        M[(long long)k * Cc + j] = M[(long long)p * Cc + j];
        M[(long long)p * Cc + j] = t;
      }
    }
    __syncthreads();
    W2 piv = M[(long long)k * Cc + k];
    if (!(pmag > 0.0) || !isfinite(pmag)) {
      if (tid == 0) atomicMax(info, 1);
      piv = CTraits<W2>::make(1, 0);
    }
    const W2 inv = cinvT(piv);
    for (int j = k + 1 + tid; j < Cc; j += nt) {
      W2* t = &M[(long long)k * Cc + j];
      *t = cmul(*t, inv);
    }
    for (int i = tid; i < R; i += nt) colk[i] = (i == k) ? zero : M[(long long)i * Cc + k];
    __syncthreads();
    for (int j = k + 1 + tid; j < Cc; j += nt) {
      const W2 rk = M[(long long)k * Cc + j];
      if (rk.x == 0 && rk.y == 0) continue;
      for (int i = 0; i < R; ++i) {
        if (i == k) continue;
        W2* t = &M[(long long)i * Cc + j];
        *t = cfms(colk[i], rk, *t);
      }
    }
    for (int i = tid; i < R; i += nt) M[(long long)i * Cc + k] = (i == k) ? CTraits<W2>::make(1, 0) : zero;
    __syncthreads();
  }
  frontStore(cd, f, factor, schur);
}

// ---- small fronts in registers
// Fronts of at most kMidRows rows (the leaves and the first heights above
// them): one workgroup per front, the R x R front held in registers
// (thread (ti, tj) of a 32 x CT grid owns rows ti + 32a and columns
// tj + CT c), no workspace.  Gauss-Jordan exchange steps in place:
// pivot (p, k) exchanges equation p and unknown k,
//   M[p][k] <- 1/M[p][k],  M[p][j] <- -M[p][j]/M[p][k],
//   M[i][k] <- M[i][k]/M[p][k],  M[i][j] <- M[i][j] - M[i][k] M[p][j]/M[p][k],
// for k = 0..e-1 with p the largest-modulus row among the fully-summed rows
// not yet exchanged (partial pivoting, as the workspace schedules; rows are
// not moved, the pivot order is kept instead).  After the e steps, with p_k
// the pivot row of column k,
//   M[p_k][l]     = (F11^{-1})[k][p_l],     M[p_k][e+t] = -(F11^{-1} F12)[k][t],
//   M[e+t][l]     = (F21 F11^{-1})[t][p_l], M[e+t][e+u] = S[t][u],
// which are W, K and S in the workspace schedules' layouts.  Column k's
// owners are the 32 lanes of one half wavefront: they pick the pivot with
// lane shuffles and publish it with the column; one workgroup barrier per
// step, the row going round by lane shuffles.
// Shapes: 96 rows as 3 rows x 6 columns a thread (512 threads; FP32
// arithmetic) or 3 x 3 (1024 threads; FP64), 128 rows as 4 x 8 (512
// threads); the thread grid is 32 rows of threads by ROWS / PC columns.
constexpr int kSmallRows = 96;
constexpr int kMidRows = 128;
constexpr int kBig32Rows = 192;  // FP32 arithmetic only

// Lane exchanges without the LDS crossbar (ds_bpermute): a value of lane
// `src` (wave-uniform) read by every lane (v_readlane per 32-bit word), and
// a DPP move within rows of 16 lanes.
template <typename T>
__device__ __forceinline__ T laneRead(T v, int src) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  int w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = __builtin_amdgcn_readlane(w[i], src);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

template <int CTRL, typename T>
__device__ __forceinline__ T dppMove(T v) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  int w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = __builtin_amdgcn_mov_dpp(w[i], CTRL, 0xF, 0xF, false);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// (modulus, row) pairs: the larger modulus wins, the lower row on ties
template <typename MT>
__device__ __forceinline__ void pivotMerge(MT& best, int& bi, MT om, int oi) {
  if (om > best || (om == best && oi < bi)) {
    best = om;
    bi = oi;
  }
}

// The pivot of a column held by the 32 lanes of one half wavefront (`half`:
// 0 for lanes 0-31, 1 for 32-63, uniform): butterflies within each row of 16
// lanes (DPP quad permutations [1,0,3,2] and [2,3,0,1], half-row and row
// mirrors), then the two rows' winners by lane reads.  The merge is a total
// order, so the winner is the same as any other reduction order's.
template <typename MT>
__device__ __forceinline__ void pivotHalfWave(MT& best, int& bi, int half) {
  pivotMerge(best, bi, dppMove<0xB1>(best), dppMove<0xB1>(bi));
  pivotMerge(best, bi, dppMove<0x4E>(best), dppMove<0x4E>(bi));
  pivotMerge(best, bi, dppMove<0x141>(best), dppMove<0x141>(bi));
  pivotMerge(best, bi, dppMove<0x140>(best), dppMove<0x140>(bi));
  MT b0 = laneRead(best, 32 * half);
  int i0 = laneRead(bi, 32 * half);
  pivotMerge(b0, i0, laneRead(best, 32 * half + 16), laneRead(bi, 32 * half + 16));
  best = b0;
  bi = i0;
}

// The same over the whole wavefront: the four rows' winners by lane reads.
template <typename MT>
__device__ __forceinline__ void pivotWave(MT& best, int& bi) {
  pivotMerge(best, bi, dppMove<0xB1>(best), dppMove<0xB1>(bi));
  pivotMerge(best, bi, dppMove<0x4E>(best), dppMove<0x4E>(bi));
  pivotMerge(best, bi, dppMove<0x141>(best), dppMove<0x141>(bi));
  pivotMerge(best, bi, dppMove<0x140>(best), dppMove<0x140>(bi));
  MT b0 = laneRead(best, 0);
  int i0 = laneRead(bi, 0);
#pragma unroll
  for (int r = 16; r < 64; r += 16) pivotMerge(b0, i0, laneRead(best, r), laneRead(bi, r));
  best = b0;
  bi = i0;
}

template <typename W2, typename O2, int PR, int PC>
__global__ __launch_bounds__(32 * (32 * PR / PC)) void k_front_small(
    const CellDev* __restrict__ cells, const ShapeDev* __restrict__ shapes, const FrontTask* __restrict__ fts,
    const double2* __restrict__ coef, O2* __restrict__ factor, O2* __restrict__ schur, int* __restrict__ info) {
  constexpr int RT = 32, ROWS = RT * PR, CT = ROWS / PC, NT = RT * CT;
  static_assert(CT * PC == ROWS && CT <= 32, "thread grid");
  static_assert(RT == 32, "a column group is one half wavefront");
  __shared__ int rowCnt[ROWS];
  __shared__ int rowCol[ROWS][5];  // a row's original entries: column
  __shared__ int rowAsm[ROWS][5];  // ... and assembly-list index
  __shared__ int inv1[ROWS], inv2[ROWS];  // front row -> child boundary index
  __shared__ int pivOf[ROWS], stepOf[ROWS];
  __shared__ W2 colK[2][ROWS];
  __shared__ double sMag[2];
  __shared__ int sP[2];
  const FrontTask ft = fts[blockIdx.x];
  const CellDev& cd = cells[ft.cell];
  const ShapeDev& sh = shapes[cd.shape];
  const NdNodeDev nd = sh.nodes[ft.node];
  const int e = nd.e, b = nd.b, R = e + b;
  const int N = cd.nx * cd.ny;
  const int tid = threadIdx.x, ti = tid % RT, tj = tid / RT;
  // ---- assembly: original entries, then child 1's and child 2's Schur
  // complements, each position in that order (as the workspace schedules)
  for (int i = tid; i < ROWS; i += NT) {
    rowCnt[i] = 0;
    inv1[i] = -1;
    inv2[i] = -1;
  }
  __syncthreads();
  for (int a = tid; a < nd.asmCnt; a += NT) {
    const AsmEntry en = sh.asmList[nd.asmOff + a];
    const int s = atomicAdd(&rowCnt[en.fr], 1);  // at most 5 entries a row, one per column
    rowCol[en.fr][s] = en.fc;
    rowAsm[en.fr][s] = a;
  }
  int bc1 = 0, bc2 = 0;
  const O2* S1 = nullptr;
  const O2* S2 = nullptr;
  if (nd.c1 >= 0) {
    const NdNodeDev cn = sh.nodes[nd.c1];
    bc1 = cn.b;
    S1 = schur + cd.schurOff + cn.sOff;
    for (int t = tid; t < bc1; t += NT) inv1[sh.up[cn.upOff + t]] = t;
  }
  if (nd.c2 >= 0) {
    const NdNodeDev cn = sh.nodes[nd.c2];
    bc2 = cn.b;
    S2 = schur + cd.schurOff + cn.sOff;
    for (int t = tid; t < bc2; t += NT) inv2[sh.up[cn.upOff + t]] = t;
  }
  __syncthreads();
  W2 M[PR][PC];
#pragma unroll
  for (int a = 0; a < PR; ++a) {
    const int i = ti + RT * a;
    const int cnt = i < R ? rowCnt[i] : 0;
    const int r1 = i < R ? inv1[i] : -1, r2 = i < R ? inv2[i] : -1;
#pragma unroll
    for (int c = 0; c < PC; ++c) {
      const int j = tj + CT * c;
      W2 v = czero<W2>();
      if (i < R && j < R) {
        for (int s2 = 0; s2 < cnt; ++s2)
          if (rowCol[i][s2] == j) {
            const AsmEntry en = sh.asmList[nd.asmOff + rowAsm[i][s2]];
            const double2 cf = coef[cd.coefOff + (long long)en.dir * N + en.dof];
            v = cadd(v, CTraits<W2>::make((typename CTraits<W2>::R)cf.x, (typename CTraits<W2>::R)cf.y));
          }
        const int q1 = inv1[j], q2 = inv2[j];
        if (r1 >= 0 && q1 >= 0) v = cadd(v, ccast<W2>(S1[(long long)r1 * bc1 + q1]));
        if (r2 >= 0 && q2 >= 0) v = cadd(v, ccast<W2>(S2[(long long)r2 * bc2 + q2]));
      }
      M[a][c] = v;
    }
  }
  // ---- the e exchange steps (one workgroup barrier each)
  // Column k's owners (tj == k % CT: the 32 lanes of one half wavefront)
  // publish it and its pivot row p in LDS; row p's values reach the other
  // threads of their column group by lane reads within the half wavefront
  // (thread (p % RT, tj) owns them), so no second barrier.
  unsigned usedMask = 0;  // this thread's rows already exchanged
  for (int k = 0; k < e; ++k) {
    const int kb = k / CT, kc = k % CT;
    W2* ck = colK[k & 1];
    if (tj == kc) {
      // (the modulus in the arithmetic's precision: the same values as
      // their conversions to double, so the same pivots)
      using MT = typename CTraits<W2>::R;
      MT best = -1;
      int bi = e;
#pragma unroll
      for (int c = 0; c < PC; ++c)
        if (c == kb) {
#pragma unroll
          for (int a = 0; a < PR; ++a) {
            const int i = ti + RT * a;
            if (i < R) ck[i] = M[a][c];
            if (i < e && !((usedMask >> a) & 1u)) {
              const MT m = M[a][c].x * M[a][c].x + M[a][c].y * M[a][c].y;
              if (m > best) {
                best = m;
                bi = i;
              }
            }
          }
        }
      // largest modulus over the 32 lanes (lowest row on ties); this half
      // wavefront is tj's, and tj's parity is the half
      pivotHalfWave(best, bi, kc & 1);
      if (ti == 0) {
        sMag[k & 1] = (double)best;
        sP[k & 1] = bi;
      }
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(sP[k & 1]);  // (uniform: the lane reads below)
    const double pmag = sMag[k & 1];
    const bool bad = !(pmag > 0.0) || !isfinite(pmag) || p >= e;
    const int pp = __builtin_amdgcn_readfirstlane(bad ? 0 : p);
    const int pa = pp / RT, pr = pp % RT;
    W2 pv = ck[pp];
    if (bad) {
      if (tid == 0) atomicMax(info, 1);
      pv = CTraits<W2>::make(1, 0);
    }
    const W2 inv = cinvT(pv);
    if (ti == pr) usedMask |= 1u << pa;
    if (tid == 0) pivOf[k] = pp;
    W2 ca[PR];
#pragma unroll
    for (int a = 0; a < PR; ++a) {
      const int i = ti + RT * a;
      ca[a] = ck[i < R ? i : 0];
    }
    const bool upper = (tid & 32) != 0;  // this lane's half wavefront: thread (pr, tj) is lane pr of it
#pragma unroll
    for (int c = 0; c < PC; ++c) {
      // row p's value in this column (before this step's update)
      W2 mine = M[0][c];
#pragma unroll
      for (int a = 1; a < PR; ++a)
        if (a == pa) mine = M[a][c];
      const W2 rowLo = laneRead(mine, pr), rowHi = laneRead(mine, pr + 32);
      const W2 rowv = upper ? rowHi : rowLo;
      const W2 rs = cmul(rowv, inv);
#pragma unroll
      for (int a = 0; a < PR; ++a) M[a][c] = cfms(ca[a], rs, M[a][c]);
      if (c == kb && tj == kc) {
#pragma unroll
        for (int a = 0; a < PR; ++a) M[a][c] = cmul(ca[a], inv);
      }
      if (ti == pr) {
        const W2 rowNew = (c == kb && tj == kc) ? inv : cneg(rs);
#pragma unroll
        for (int a = 0; a < PR; ++a)
          if (a == pa) M[a][c] = rowNew;
      }
    }
  }
  __syncthreads();
  for (int k = tid; k < e; k += NT) stepOf[pivOf[k]] = k;
  __syncthreads();
  // ---- W, K and S
  O2* Wp = factor + cd.factorOff + nd.fwdOff;
  O2* Kp = factor + cd.factorOff + nd.bwdOff;
  O2* Sp = schur + cd.schurOff + nd.sOff;
#pragma unroll
  for (int a = 0; a < PR; ++a) {
    const int i = ti + RT * a;
    if (i >= R) continue;
    const int ki = i < e ? stepOf[i] : 0;
#pragma unroll
    for (int c = 0; c < PC; ++c) {
      const int j = tj + CT * c;
      if (j >= R) continue;
      const O2 v = ccast<O2>(M[a][c]);
      if (i < e) {
        if (j < e)
          Kp[(long long)pivOf[j] * e + ki] = v;
        else
          Kp[(long long)j * e + ki] = v;
      } else if (j < e) {
        Wp[(long long)pivOf[j] * b + (i - e)] = v;
      } else {
        Sp[(long long)(i - e) * b + (j - e)] = v;
      }
    }
  }
}

// ---- blocked schedule for the large fronts
// Per front, after the front itself (R x Cc) the workspace holds T (R x
// kPanel, row-major) and B (kPanel x Cc, row-major).
constexpr int kPanel = 16;

template <typename W2, typename O2>
__global__ __launch_bounds__(256) void k_front_assemble(const CellDev* __restrict__ cells,
                                                        const ShapeDev* __restrict__ shapes,
                                                        const FrontTask* __restrict__ fts,
                                                        const double2* __restrict__ coef,
                                                        const O2* __restrict__ schur, W2* __restrict__ ws) {
  const FrontTask ft = fts[blockIdx.x];
  const CellDev& cd = cells[ft.cell];
  frontAssemble(cd, shapes[cd.shape], frontView(cells, shapes, ft, ws), coef, schur);
}

template <typename W2, typename O2>
__global__ __launch_bounds__(256) void k_front_store(const CellDev* __restrict__ cells,
                                                     const ShapeDev* __restrict__ shapes,
                                                     const FrontTask* __restrict__ fts, O2* __restrict__ factor,
                                                     O2* __restrict__ schur, W2* __restrict__ ws) {
  const FrontTask ft = fts[blockIdx.x];
  frontStore(cells[ft.cell], frontView(cells, shapes, ft, ws), factor, schur);
}

// Pivots k0 .. k0+nbk-1 of one front.  Gauss-Jordan with partial pivoting
// on the panel columns only, with the row swaps applied to whole rows (from
// column k0 on) and the eliminations recorded in T: T column c starts as the
// unit vector of pivot row k0+c when that pivot is reached and takes every
// later scaling, swap and elimination, so that at the end, in the swapped
// row order, T = [A11^{-1}; -A21 A11^{-1}] (A11 = the pivot rows' panel
// block, A21 = the other rows').  The rest of the front is then
//   pivot rows: new = T B,  other rows: new = old + T B,
// B = old pivot rows (columns after the panel): the batched update below.
// (Not old + (T - E) B with E = 1 at (pivot row c, c), which is the same
// in exact arithmetic: T - E rounds A11^{-1}'s diagonal against 1, which
// lost most FP32 digits where |A11^{-1}| << 1.)  B is copied here, before
// that update changes the pivot rows.
template <typename W2>
__global__ __launch_bounds__(256) void k_front_panel(const CellDev* __restrict__ cells,
                                                     const ShapeDev* __restrict__ shapes,
                                                     const FrontTask* __restrict__ fts, W2* __restrict__ ws,
                                                     int k0, int* __restrict__ info) {
  extern __shared__ double2 lswDynLds[];
  W2* colk = reinterpret_cast<W2*>(lswDynLds);
  __shared__ double redv[256];
  __shared__ int redi[256];
  const FrontTask ft = fts[blockIdx.x];
  const FrontView<W2> f = frontView(cells, shapes, ft, ws);
  const int e = f.e, R = f.R, Cc = f.Cc;
  const int nbk = min(kPanel, e - k0);
  if (nbk <= 0) return;
  W2* M = f.M;
  W2* T = M + (long long)R * Cc;
  W2* B = T + (long long)R * kPanel;
  const int tid = threadIdx.x, nt = blockDim.x;
  const W2 zero = czero<W2>();
  const int pend = k0 + nbk;  // panel columns [k0, pend)
  for (int idx = tid; idx < R * kPanel; idx += nt) T[idx] = zero;
  __syncthreads();
  for (int kk = 0; kk < nbk; ++kk) {
    const int k = k0 + kk;
    double pmag;
    const int p = pivotRow(M, Cc, k, e, redv, redi, pmag);
    if (p != k) {
      for (int j = k0 + tid; j < Cc; j += nt) {
        const W2 t = M[(long long)k * Cc + j];
        M[(long long)k * Cc + j] = M[(long long)p * Cc + j];
        M[(long long)p * Cc + j] = t;
      }
      for (int c = tid; c < kk; c += nt) {
        const W2 t = T[k * kPanel + c];
        T[k * kPanel + c] = T[p * kPanel + c];
        T[p * kPanel + c] = t;
      }
    }
    __syncthreads();
    if (tid == 0) T[k * kPanel + kk] = CTraits<W2>::make(1, 0);
    W2 piv = M[(long long)k * Cc + k];
    if (!(pmag > 0.0) || !isfinite(pmag)) {
      if (tid == 0) atomicMax(info, 1);
      piv = CTraits<W2>::make(1, 0);
    }
    const W2 inv = cinvT(piv);
    __syncthreads();
    // scale the pivot row (panel columns after k, and T)
    for (int j = tid; j < (pend - k - 1) + (kk + 1); j += nt) {
      W2* t = j < pend - k - 1 ? &M[(long long)k * Cc + k + 1 + j] : &T[k * kPanel + (j - (pend - k - 1))];
      *t = cmul(*t, inv);
    }
    for (int i = tid; i < R; i += nt) colk[i] = (i == k) ? zero : M[(long long)i * Cc + k];
    __syncthreads();
    // eliminate column k from every other row (panel columns after k, and T)
    const int w = (pend - k - 1) + (kk + 1);
    for (long long idx = tid; idx < (long long)R * w; idx += nt) {
      const int i = (int)(idx / w), j = (int)(idx - (long long)i * w);
      if (i == k) continue;
      const W2 fi = colk[i];
      if (fi.x == 0 && fi.y == 0) continue;
      if (j < pend - k - 1) {
        W2* t = &M[(long long)i * Cc + k + 1 + j];
        *t = cfms(fi, M[(long long)k * Cc + k + 1 + j], *t);
      } else {
        const int c = j - (pend - k - 1);
        W2* t = &T[i * kPanel + c];
        *t = cfms(fi, T[k * kPanel + c], *t);
      }
    }
    for (int i = tid; i < R; i += nt) M[(long long)i * Cc + k] = (i == k) ? CTraits<W2>::make(1, 0) : zero;
    __syncthreads();
  }
  // B = the pivot rows from the first column after the panel
  const int cw = Cc - pend;
  for (long long idx = tid; idx < (long long)nbk * cw; idx += nt) {
    const int c = (int)(idx / cw), j = (int)(idx - (long long)c * cw);
    B[(long long)c * Cc + j] = M[(long long)(k0 + c) * Cc + pend + j];
  }
}

// The same panel with its columns, T and the elimination column held in
// LDS (fronts of up to kPanelLdsRows rows): the 16 pivot steps then run on
// LDS instead of chains of L2 round trips, and the pivot search reduces
// within each wavefront by DPP moves and lane reads (one barrier instead of
// a 256-wide LDS tree).  The row swaps of the columns after the panel are recorded and
// applied at the end, column by column (each thread owns columns and
// replays the swaps in order), together with the copy of B.  Same
// arithmetic in the same order per entry as k_front_panel, and the same
// pivots (largest modulus, lowest row on ties, an order-free choice).
// Up to 280 rows in FP64, 576 in FP32 (157 KB of the 160 KB LDS).
constexpr int kPanelLdsRows = 280;
constexpr int kPanelLdsRows32 = 576;
constexpr int kPanelLd = kPanel + 1;  // padded LDS row (bank spread)

__host__ __device__ constexpr size_t panelLdsBytes(int rows, size_t es) {
  return (size_t)rows * (2 * kPanelLd + 1) * es;
}

template <typename W2>
__global__ __launch_bounds__(256) void k_front_panel_lds(const CellDev* __restrict__ cells,
                                                         const ShapeDev* __restrict__ shapes,
                                                         const FrontTask* __restrict__ fts,
                                                         W2* __restrict__ ws, int k0, int* __restrict__ info) {
  using MT = typename CTraits<W2>::R;  // pivot moduli (the values their conversions to double have)
  extern __shared__ double2 lswDynLds[];
  __shared__ MT wbest[4];
  __shared__ int wrow[4];
  __shared__ int perm[kPanel];
  const FrontTask ft = fts[blockIdx.x];
  const FrontView<W2> f = frontView(cells, shapes, ft, ws);
  const int e = f.e, R = f.R, Cc = f.Cc;
  const int nbk = min(kPanel, e - k0);
  if (nbk <= 0) return;
  W2* sP = reinterpret_cast<W2*>(lswDynLds);  // R x kPanelLd: panel columns k0 .. k0+nbk-1
  W2* sT = sP + R * kPanelLd;                  // R x kPanelLd
  W2* sC = sT + R * kPanelLd;                  // R: the column being eliminated
  W2* M = f.M;
  W2* T = M + (long long)R * Cc;
  W2* B = T + (long long)R * kPanel;
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
  const W2 zero = czero<W2>(), one = CTraits<W2>::make(1, 0);
  const int pend = k0 + nbk;
  for (int idx = tid; idx < R * kPanel; idx += nt) {
    const int i = idx / kPanel, c = idx - i * kPanel;
    sP[i * kPanelLd + c] = c < nbk ? M[(long long)i * Cc + k0 + c] : zero;
    sT[i * kPanelLd + c] = zero;
  }
  __syncthreads();
  for (int kk = 0; kk < nbk; ++kk) {
    const int k = k0 + kk;
    MT best = -1;
    int bi = e;
    for (int i = k + tid; i < e; i += nt) {
      const W2 v = sP[i * kPanelLd + kk];
      const MT a = v.x * v.x + v.y * v.y;
      if (a > best) {
        best = a;
        bi = i;
      }
    }
    pivotWave(best, bi);
    if (lane == 0) {
      wbest[wave] = best;
      wrow[wave] = bi;
    }
    __syncthreads();
    best = wbest[0];
    bi = wrow[0];
    for (int w = 1; w < nt / 64; ++w)
      if (wbest[w] > best || (wbest[w] == best && wrow[w] < bi)) {
        best = wbest[w];
        bi = wrow[w];
      }
    const int p = bi;
    const double pmag = (double)best;
    if (tid == 0) perm[kk] = p;
    // columns before kk are unit columns of rows above k: zero in rows k, p
    if (p != k) {
      for (int j = tid; j < (nbk - kk) + kk; j += nt) {
        W2* a = j < nbk - kk ? &sP[k * kPanelLd + kk + j] : &sT[k * kPanelLd + (j - (nbk - kk))];
        W2* b = j < nbk - kk ? &sP[p * kPanelLd + kk + j] : &sT[p * kPanelLd + (j - (nbk - kk))];
        const W2 t = *a;
        *a = *b;
        *b = t;
      }
    }
    __syncthreads();
    if (tid == 0) sT[k * kPanelLd + kk] = one;
    W2 piv = sP[k * kPanelLd + kk];
    if (!(pmag > 0.0) || !isfinite(pmag)) {
      if (tid == 0) atomicMax(info, 1);
      piv = one;
    }
    const W2 inv = cinvT(piv);
    __syncthreads();
    for (int j = tid; j < (nbk - kk - 1) + (kk + 1); j += nt) {
      W2* t = j < nbk - kk - 1 ? &sP[k * kPanelLd + kk + 1 + j] : &sT[k * kPanelLd + (j - (nbk - kk - 1))];
      *t = cmul(*t, inv);
    }
    for (int i = tid; i < R; i += nt) sC[i] = (i == k) ? zero : sP[i * kPanelLd + kk];
    __syncthreads();
    const int w = (nbk - kk - 1) + (kk + 1);
    for (int idx = tid; idx < R * w; idx += nt) {
      const int i = idx / w, j = idx - i * w;
      if (i == k) continue;
      const W2 fi = sC[i];
      if (fi.x == 0 && fi.y == 0) continue;
      if (j < nbk - kk - 1) {
        W2* t = &sP[i * kPanelLd + kk + 1 + j];
        *t = cfms(fi, sP[k * kPanelLd + kk + 1 + j], *t);
      } else {
        const int c = j - (nbk - kk - 1);
        W2* t = &sT[i * kPanelLd + c];
        *t = cfms(fi, sT[k * kPanelLd + c], *t);
      }
    }
    for (int i = tid; i < R; i += nt) sP[i * kPanelLd + kk] = (i == k) ? one : zero;
    __syncthreads();
  }
  // panel columns and T back to the workspace
  for (int idx = tid; idx < R * kPanel; idx += nt) {
    const int i = idx / kPanel, c = idx - i * kPanel;
    if (c < nbk) M[(long long)i * Cc + k0 + c] = sP[i * kPanelLd + c];
    T[i * kPanel + c] = sT[i * kPanelLd + c];
  }
  // the recorded swaps on the columns after the panel, then B
  for (int j = pend + tid; j < Cc; j += nt) {
    for (int kk = 0; kk < nbk; ++kk) {
      const int k = k0 + kk, p = perm[kk];
      if (p != k) {
        const W2 t = M[(long long)k * Cc + j];
        M[(long long)k * Cc + j] = M[(long long)p * Cc + j];
        M[(long long)p * Cc + j] = t;
      }
    }
    for (int c = 0; c < nbk; ++c) B[(long long)c * Cc + (j - pend)] = M[(long long)(k0 + c) * Cc + j];
  }
}

// new = T B (the panel's pivot rows) or old + T B (the other rows) over rows
// [0, R) and columns [pend, Cc) of every front: 64 x 64 output tiles, 4 x 4
// per thread, the panel's T and B slices
// staged in LDS.  blockIdx.y = front, blockIdx.x = tile.  (FP64 fronts; kept
// for comparison with the matrix-core kernel.)
__global__ __launch_bounds__(256) void k_front_trail(const CellDev* __restrict__ cells,
                                                     const ShapeDev* __restrict__ shapes,
                                                     const FrontTask* __restrict__ fts, double2* __restrict__ ws,
                                                     int k0) {
  __shared__ double2 sT[64][kPanel + 1];
  __shared__ double2 sB[kPanel][64 + 1];
  const FrontTask ft = fts[blockIdx.y];
  const FrontView<double2> f = frontView(cells, shapes, ft, ws);
  const int e = f.e, R = f.R, Cc = f.Cc;
  const int nbk = min(kPanel, e - k0);
  if (nbk <= 0) return;
  const int pend = k0 + nbk, cw = Cc - pend;
  const int tilesC = (cw + 63) / 64, tilesR = (R + 63) / 64;
  const int tile = blockIdx.x;
  if (tile >= tilesC * tilesR) return;
  const int r0 = (tile / tilesC) * 64, c0 = (tile % tilesC) * 64;
  double2* M = f.M;
  const double2* T = M + (long long)R * Cc;
  const double2* B = T + (long long)R * kPanel;
  const int tid = threadIdx.x;
  const double2 zero = make_double2(0.0, 0.0);
  for (int idx = tid; idx < 64 * kPanel; idx += 256) {
    const int i = idx / kPanel, c = idx - i * kPanel;
    sT[i][c] = (r0 + i < R && c < nbk) ? T[(r0 + i) * kPanel + c] : zero;
  }
  for (int idx = tid; idx < kPanel * 64; idx += 256) {
    const int c = idx / 64, j = idx - c * 64;
    sB[c][j] = (c < nbk && c0 + j < cw) ? B[(long long)c * Cc + c0 + j] : zero;
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  double2 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[a][q] = zero;
  for (int c = 0; c < nbk; ++c) {
    double2 tv[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) tv[a] = sT[ty * 4 + a][c];
#pragma unroll
    for (int q = 0; q < 4; ++q) bv[q] = sB[c][tx * 4 + q];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[a][q] = cfma(tv[a], bv[q], acc[a][q]);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = r0 + ty * 4 + a;
    if (i >= R) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = c0 + tx * 4 + q;
      if (j >= cw) continue;
      double2* t = &M[(long long)i * Cc + pend + j];
      *t = (i >= k0 && i < pend) ? acc[a][q] : cadd(*t, acc[a][q]);
    }
  }
}

// The same update on the matrix cores: each wavefront owns a 32 x 32
// quadrant of the 64 x 64 tile as 2 x 2 blocks of 16 x 16, and every complex
// product-accumulate is four real 16x16x4 matrix-core products (Cr += Tr Br -
// Ti Bi, Ci += Tr Bi + Ti Br) over k-steps of 4 panel columns:
// v_mfma_f64_16x16x4f64 on FP64 fronts, v_mfma_f32_16x16x4f32 (FP32
// accumulation) on FP32 fronts.  Operand layout of both (checked on gfx950):
// lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; the f64 result
// register r of lane l holds D[l / 16 + 4 r][l % 16], the f32 one
// D[4 (l / 16) + r][l % 16].
typedef double lswD4 __attribute__((ext_vector_type(4)));
typedef float lswS4 __attribute__((ext_vector_type(4)));

template <typename W2>
struct TrailAcc;
template <>
struct TrailAcc<double2> {
  using V = lswD4;
  __device__ static V mma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct TrailAcc<float2> {
  using V = lswS4;
  __device__ static V mma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
};

template <typename W2>
__global__ __launch_bounds__(256) void k_front_trail_mfma(const CellDev* __restrict__ cells,
                                                          const ShapeDev* __restrict__ shapes,
                                                          const FrontTask* __restrict__ fts,
                                                          W2* __restrict__ ws, int k0) {
  using A = TrailAcc<W2>;
  using V = typename A::V;
  __shared__ W2 sT[64][kPanel + 1];
  __shared__ W2 sB[kPanel][64 + 1];
  const FrontTask ft = fts[blockIdx.y];
  const FrontView<W2> f = frontView(cells, shapes, ft, ws);
  const int e = f.e, R = f.R, Cc = f.Cc;
  const int nbk = min(kPanel, e - k0);
  if (nbk <= 0) return;
  const int pend = k0 + nbk, cw = Cc - pend;
  const int tilesC = (cw + 63) / 64, tilesR = (R + 63) / 64;
  const int tile = blockIdx.x;
  if (tile >= tilesC * tilesR) return;
  const int r0 = (tile / tilesC) * 64, c0 = (tile % tilesC) * 64;
  W2* M = f.M;
  const W2* T = M + (long long)R * Cc;
  const W2* B = T + (long long)R * kPanel;
  const int tid = threadIdx.x;
  const W2 zero = czero<W2>();
  for (int idx = tid; idx < 64 * kPanel; idx += 256) {
    const int i = idx / kPanel, c = idx - i * kPanel;
    sT[i][c] = (r0 + i < R && c < nbk) ? T[(r0 + i) * kPanel + c] : zero;
  }
  for (int idx = tid; idx < kPanel * 64; idx += 256) {
    const int c = idx / 64, j = idx - c * 64;
    sB[c][j] = (c < nbk && c0 + j < cw) ? B[(long long)c * Cc + c0 + j] : zero;
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int qr = (wave >> 1) * 32, qc = (wave & 1) * 32;  // the wavefront's quadrant
  const int li = lane & 15, lk = lane >> 4;
  V cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = V{0, 0, 0, 0};
      ci[a][b] = V{0, 0, 0, 0};
    }
  for (int kk = 0; kk < nbk; kk += 4) {  // columns past nbk are zero in the staging
    W2 ta[2], tb[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) ta[a] = sT[qr + a * 16 + li][kk + lk];
#pragma unroll
    for (int b = 0; b < 2; ++b) tb[b] = sB[kk + lk][qc + b * 16 + li];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        cr[a][b] = A::mma(ta[a].x, tb[b].x, cr[a][b]);
        cr[a][b] = A::mma(-ta[a].y, tb[b].y, cr[a][b]);
        ci[a][b] = A::mma(ta[a].x, tb[b].y, ci[a][b]);
        ci[a][b] = A::mma(ta[a].y, tb[b].x, ci[a][b]);
      }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = c0 + qc + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = r0 + qr + a * 16 + A::row(lk, r);
        if (i < R && j < cw) {
          W2* t = &M[(long long)i * Cc + pend + j];
          const W2 v = CTraits<W2>::make(cr[a][b][r], ci[a][b][r]);
          *t = (i >= k0 && i < pend) ? v : cadd(*t, v);
        }
      }
    }
}

__global__ void k_to_f32(const double2* __restrict__ a, float2* __restrict__ o, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    o[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// --------------------------------------------------------------- host side
FactorSet::~FactorSet() {
  mArena.release();
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
      auto t = sharedTree(c.nx, c.ny);
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
    d.nodes = mArena.put(t.deviceNodes());
    d.E = mArena.put(t.E);
    d.B = mArena.put(t.B);
    d.cmap = mArena.put(t.cmap);
    d.up = mArena.put(t.up);
    d.asmList = mArena.put(t.asmList);
    std::vector<SolveItem> fi, bi;
    for (int h = 0; h <= t.maxHeight; ++h) {
      d.fwdStart[h] = (int)fi.size();
      d.fwdCount[h] = (int)t.fwdItems[h].size();
      fi.insert(fi.end(), t.fwdItems[h].begin(), t.fwdItems[h].end());
      d.bwdStart[h] = (int)bi.size();
      d.bwdCount[h] = (int)t.bwdItems[h].size();
      bi.insert(bi.end(), t.bwdItems[h].begin(), t.bwdItems[h].end());
    }
    d.fwdItems = mArena.put(fi);
    d.bwdItems = mArena.put(bi);
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
    dSolveNodes = mArena.put(flat);
    mSolveNodes = flat;
    dPoolE = mArena.put(pe);
    dPoolB = mArena.put(pb);
    dPoolMap = mArena.put(pm);
  }
  dShapes = mArena.put(shapes);
  dCells = mArena.put(mCells);
  mArena.flush();
  // stencils
  LSW_HIP(devMalloc(&dCoef, co * sizeof(double2)));
  for (size_t i = 0; i < mCells.size(); ++i)
    if (!mCoefHost[i].empty())
      LSW_HIP(uploadSync(dCoef + mCells[i].coefOff, mCoefHost[i].data(),
                        mCoefHost[i].size() * sizeof(double2)));
  if (mo > 0) {
    // every cell's squared slowness in one upload
    LSW_HIP(devMalloc(&dM, mo * sizeof(double)));
    {
      std::vector<double> all((size_t)mo);
      for (size_t i = 0; i < mCells.size(); ++i)
        if (!mM[i].empty()) std::copy(mM[i].begin(), mM[i].end(), all.begin() + mCells[i].mOff);
      LSW_HIP(uploadSync(dM, all.data(), (size_t)mo * sizeof(double)));
    }
    // the PML evaluation range on the host (no device round trip), then the
    // stencils on the stream: the factorisation reads them in order
    for (const CellDev& c : mCells) {
      if (c.mOff < 0) continue;
      Geo2 g;
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          g.intDom[a][b] = c.intDom[a][b];
          g.compDom[a][b] = c.compDom[a][b];
        }
      LSW_REQUIRE(gridInsidePml(g, c.x0, c.y0, c.hx, c.hy, c.nx, c.ny), LSW_ERR_ARG,
                  "PML: Evaluation point out of range!");
    }
    dim3 grid((unsigned)std::min<long long>((mMaxDofs + 255) / 256, 4096), (unsigned)mCells.size());
    hipLaunchKernelGGL(k_assemble, grid, dim3(256), 0, s, dCells, dM, dCoef);
    LSW_HIP(hipGetLastError());
  }
}

namespace {
// Heights whose largest front has at least this many rows take the blocked
// schedule.  Default 1 (every height): on the 1600^2 lens workload the
// factorisation took 0.39 s with one workgroup per front everywhere, 0.20 s
// with the blocked schedule from 192 rows up and 0.149 s with it on every
// height.  LSW_FACTOR_BLOCKED_ROWS overrides it for tuning.
// Heights whose fronts all have at most kMidRows rows take k_front_small
// (LSW_FACTOR_SMALL=0: the workspace schedules, for comparison).
bool smallFronts() {
  static const bool v = [] {
    const char* e = std::getenv("LSW_FACTOR_SMALL");
    return !(e && std::atoi(e) == 0);
  }();
  return v;
}

// With FP64 fronts for an FP32-factor set (LSW_FACTOR_FP32=0), the small
// fronts (k_front_small) still run in FP32 arithmetic unless
// LSW_FACTOR_SMALL_FP32=0.
bool smallFrontsFp32() {
  static const bool v = [] {
    const char* e = std::getenv("LSW_FACTOR_SMALL_FP32");
    return !(e && std::atoi(e) == 0);
  }();
  return v;
}

int blockedMinRows() {
  static const int v = [] {
    const char* e = std::getenv("LSW_FACTOR_BLOCKED_ROWS");
    return e ? std::max(1, std::atoi(e)) : 1;
  }();
  return v;
}

// Trailing updates on the matrix cores (k_front_trail_mfma) unless
// LSW_FACTOR_MFMA=0 selects the vector-unit kernel (FP64 fronts; kept for
// comparison).
bool trailMfma() {
  static const bool v = [] {
    const char* e = std::getenv("LSW_FACTOR_MFMA");
    return !(e && std::atoi(e) == 0);
  }();
  return v;
}

// LDS-resident panels for fronts of up to kPanelLdsRows rows (FP64) or
// kPanelLdsRows32 rows (FP32) unless LSW_PANEL_LDS=0.
bool panelLds() {
  static const bool v = [] {
    const char* e = std::getenv("LSW_PANEL_LDS");
    return !(e && std::atoi(e) == 0);
  }();
  return v;
}

void launchPanel(unsigned nf, int maxR, size_t shm, hipStream_t s, const CellDev* cells, const ShapeDev* shapes,
                 const FrontTask* fts, double2* ws, int k0, int* info) {
  if (panelLds() && maxR > 0 && maxR <= kPanelLdsRows)
    hipLaunchKernelGGL(k_front_panel_lds<double2>, dim3(nf), dim3(256), panelLdsBytes(maxR, sizeof(double2)), s, cells,
                       shapes, fts, ws, k0, info);
  else
    hipLaunchKernelGGL(k_front_panel<double2>, dim3(nf), dim3(256), shm, s, cells, shapes, fts, ws, k0, info);
}

void launchPanel(unsigned nf, int maxR, size_t shm, hipStream_t s, const CellDev* cells, const ShapeDev* shapes,
                 const FrontTask* fts, float2* ws, int k0, int* info) {
  if (panelLds() && maxR > 0 && maxR <= kPanelLdsRows32)
    hipLaunchKernelGGL(k_front_panel_lds<float2>, dim3(nf), dim3(256), panelLdsBytes(maxR, sizeof(float2)), s, cells,
                       shapes, fts, ws, k0, info);
  else
    hipLaunchKernelGGL(k_front_panel<float2>, dim3(nf), dim3(256), shm, s, cells, shapes, fts, ws, k0, info);
}

void launchTrail(dim3 grid, hipStream_t s, const CellDev* cells, const ShapeDev* shapes, const FrontTask* fts,
                 double2* ws, int k0) {
  if (trailMfma())
    hipLaunchKernelGGL(k_front_trail_mfma<double2>, grid, dim3(256), 0, s, cells, shapes, fts, ws, k0);
  else
    hipLaunchKernelGGL(k_front_trail, grid, dim3(256), 0, s, cells, shapes, fts, ws, k0);
}

void launchTrail(dim3 grid, hipStream_t s, const CellDev* cells, const ShapeDev* shapes, const FrontTask* fts,
                 float2* ws, int k0) {
  hipLaunchKernelGGL(k_front_trail_mfma<float2>, grid, dim3(256), 0, s, cells, shapes, fts, ws, k0);
}
}  // namespace

// FP32-factor sets factorise in FP32 (LSW_FACTOR_FP32=0: FP64 fronts,
// FP32 copy of the factors).  M^{-1} against the FP64-factor one: 2.9e-5
// relative on a 3 x 3 board of 40 x 40 cells (FP64 fronts: 5.1e-5), and the
// same GMRES iteration counts on the 48-problem pack.  (The panel update
// once formed old + (T - E) B for the pivot rows, which cancels A11^{-1}'s
// diagonal against 1: FP32 fronts then gave 0.11 on that board.)
bool factorWorkFp32(int precision) {
  static const bool fp32Work = [] {
    const char* e = std::getenv("LSW_FACTOR_FP32");
    return !(e && std::atoi(e) == 0);
  }();
  return precision == LSW_FP32 && fp32Work;
}

void FactorSet::factorize(int precision, hipStream_t s) {
  LSW_REQUIRE(dCells != nullptr, LSW_ERR_STATE, "FactorSet: setup() not called");
  LSW_REQUIRE(precision == LSW_FP64 || precision == LSW_FP32, LSW_ERR_ARG, "unknown precision");
  if (factorWorkFp32(precision))
    factorizeAs<float2>(s, dFactor32, true);
  else
    factorizeAs<double2>(s, dFactor, precision == LSW_FP32 && smallFrontsFp32());
  mPrecision = precision;
  if (precision == LSW_FP32 && dFactor) {
    if (!dFactor32) LSW_HIP(devMalloc(&dFactor32, std::max<long long>(mFactorEntries, 1) * sizeof(float2)));
    hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, s, dFactor, dFactor32, mFactorEntries);
    LSW_HIP(hipGetLastError());
    LSW_HIP(hipStreamSynchronize(s));
    IdleFrees idle;  // the FP64 factors were only used on s
    (void)devFree(dFactor);
    dFactor = nullptr;
  }
  mFactorized = true;
}

template <typename W2>
void FactorSet::factorizeAs(hipStream_t s, W2*& factor, bool smallFp32) {
  if (!factor) LSW_HIP(devMalloc(&factor, std::max<long long>(mFactorEntries, 1) * sizeof(W2)));
  // Schur workspace for all cells
  std::vector<long long> schurOff(mCells.size());
  long long so = 0;
  for (size_t i = 0; i < mCells.size(); ++i) {
    schurOff[i] = so;
    so += mTrees[mCells[i].shape]->schurSize;
    mCells[i].schurOff = so - mTrees[mCells[i].shape]->schurSize;
  }
  LSW_HIP(uploadSync(dCells, mCells.data(), mCells.size() * sizeof(CellDev)));
  W2* dSchur = nullptr;
  LSW_HIP(devMalloc(&dSchur, std::max<long long>(so, 1) * sizeof(W2)));
  int* dInfo = nullptr;
  LSW_HIP(devMalloc(&dInfo, sizeof(int)));
  LSW_HIP(hipMemsetAsync(dInfo, 0, sizeof(int), s));
  // front batches per height, bounded by a workspace budget; a height whose
  // largest front has at least blockedMinRows() rows takes the blocked
  // schedule (panels + batched updates), which also needs T and B per front
  const long long budget = kFactorWsEntries;
  std::vector<std::vector<FrontTask>> batches;
  std::vector<long long> batchWs;
  std::vector<char> batchBlocked;
  long long maxWs = 0;
  for (int h = 0; h <= mMaxHeight; ++h) {
    int maxR = 0;
    for (size_t ci = 0; ci < mCells.size(); ++ci) {
      const NdTree& t = *mTrees[mCells[ci].shape];
      if (h > t.maxHeight) continue;
      for (int p : t.byHeight[h]) maxR = std::max(maxR, t.nodes[p].e + t.nodes[p].b);
    }
    // heights of small fronts only: one register-resident workgroup per
    // front, no workspace (state 2)
    const bool small = smallFronts() && maxR <= kSmallRows;
    const bool mid = !small && smallFronts() && maxR <= kMidRows;
    // FP32 arithmetic: fronts of up to 192 rows fit the registers of one
    // 1024-thread workgroup (6 x 6 entries a thread)
    const bool big32 = !small && !mid && smallFronts() && smallFp32 && maxR <= kBig32Rows;
    const bool blocked = maxR >= blockedMinRows();
    const char state = small ? 2 : mid ? 3 : big32 ? 4 : blocked ? 1 : 0;
    std::vector<FrontTask> cur;
    long long ws = 0;
    for (size_t ci = 0; ci < mCells.size(); ++ci) {
      const NdTree& t = *mTrees[mCells[ci].shape];
      if (h > t.maxHeight) continue;
      for (int p : t.byHeight[h]) {
        const NdNode& nd = t.nodes[p];
        const long long R = nd.e + nd.b, Cc = 2 * nd.e + nd.b;
        const long long need = (small || mid || big32) ? 0 : R * Cc + (blocked ? R * kPanel + kPanel * Cc : 0);
        if (!cur.empty() && ws + need > budget) {
          batches.push_back(cur);
          batchWs.push_back(ws);
          batchBlocked.push_back(state);
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
      batchBlocked.push_back(state);
      maxWs = std::max(maxWs, ws);
    }
    batches.push_back({});  // height separator (keeps order explicit)
    batchWs.push_back(0);
    batchBlocked.push_back(0);
  }
  auto nodeOf = [&](const FrontTask& ft) -> const NdNode& { return mTrees[mCells[ft.cell].shape]->nodes[ft.node]; };
  W2* dWs = nullptr;
  LSW_HIP(devMalloc(&dWs, std::max<long long>(maxWs, 1) * sizeof(W2)));
  // every batch's front list in one upload; the batches then run back to
  // back on the stream with no host round trip between heights
  std::vector<FrontTask> allFt;
  std::vector<size_t> ftOff(batches.size());
  for (size_t bi = 0; bi < batches.size(); ++bi) {
    ftOff[bi] = allFt.size();
    allFt.insert(allFt.end(), batches[bi].begin(), batches[bi].end());
  }
  FrontTask* dFtAll = nullptr;
  LSW_HIP(devMalloc(&dFtAll, std::max<size_t>(allFt.size(), 1) * sizeof(FrontTask)));
  LSW_HIP(uploadSync(dFtAll, allFt.data(), allFt.size() * sizeof(FrontTask)));
  for (size_t bi = 0; bi < batches.size(); ++bi) {
    const auto& bt = batches[bi];
    if (bt.empty()) continue;
    const FrontTask* dFt = dFtAll + ftOff[bi];
    const size_t shm = (size_t)mMaxFront * sizeof(W2);
    const unsigned nf = (unsigned)bt.size();
    if (batchBlocked[bi] == 2) {
      if (smallFp32)
        hipLaunchKernelGGL((k_front_small<float2, W2, 3, 6>), dim3(nf), dim3(512), 0, s, dCells, dShapes, dFt, dCoef,
                           factor, dSchur, dInfo);
      else  // (FP64: 3 x 3 a thread on 1024 threads, 116 VGPRs; 3 x 6 on 512 needed 154)
        hipLaunchKernelGGL((k_front_small<W2, W2, 3, 3>), dim3(nf), dim3(1024), 0, s, dCells, dShapes, dFt, dCoef,
                           factor, dSchur, dInfo);
    } else if (batchBlocked[bi] == 4) {
      hipLaunchKernelGGL((k_front_small<float2, W2, 6, 6>), dim3(nf), dim3(1024), 0, s, dCells, dShapes, dFt, dCoef,
                         factor, dSchur, dInfo);
    } else if (batchBlocked[bi] == 3) {
      hipLaunchKernelGGL((k_front_small<W2, W2, 4, 8>), dim3(nf), dim3(512), 0, s, dCells, dShapes, dFt, dCoef,
                         factor, dSchur, dInfo);
    } else if (!batchBlocked[bi]) {
      hipLaunchKernelGGL(k_front_factor<W2>, dim3(nf), dim3(256), shm, s, dCells, dShapes, dFt, dCoef, factor,
                         dSchur, dWs, dInfo);
    } else {
      // the workspace schedule
      auto blockedRun = [&](auto* wsT) {
        using WT = typename std::remove_pointer<decltype(wsT)>::type;
        hipLaunchKernelGGL((k_front_assemble<WT, W2>), dim3(nf), dim3(256), 0, s, dCells, dShapes, dFt, dCoef,
                           dSchur, wsT);
        int maxE = 0;
        for (const auto& ft : bt) maxE = std::max(maxE, nodeOf(ft).e);
        for (int k0 = 0; k0 < maxE; k0 += kPanel) {
          int maxR = 0;
          for (const auto& ft : bt) {
            const NdNode& nd = nodeOf(ft);
            if (nd.e > k0) maxR = std::max(maxR, nd.e + nd.b);
          }
          launchPanel(nf, maxR, shm, s, dCells, dShapes, dFt, wsT, k0, dInfo);
          long long tiles = 0;
          for (const auto& ft : bt) {
            const NdNode& nd = nodeOf(ft);
            if (nd.e <= k0) continue;
            const long long R = nd.e + nd.b, cw = 2LL * nd.e + nd.b - std::min(k0 + kPanel, nd.e);
            tiles = std::max(tiles, ((R + 63) / 64) * ((cw + 63) / 64));
          }
          if (tiles > 0) launchTrail(dim3((unsigned)tiles, nf), s, dCells, dShapes, dFt, wsT, k0);
        }
        hipLaunchKernelGGL((k_front_store<WT, W2>), dim3(nf), dim3(256), 0, s, dCells, dShapes, dFt, factor, dSchur,
                           wsT);
      };
      blockedRun(dWs);
    }
    LSW_HIP(hipGetLastError());
  }
  LSW_HIP(hipMemcpyAsync(&mInfo, dInfo, sizeof(int), hipMemcpyDeviceToHost, s));
  LSW_HIP(hipStreamSynchronize(s));
  IdleFrees idle;  // the workspace was only used on s
  (void)devFree(dFtAll);
  (void)devFree(dWs);
  (void)devFree(dSchur);
  (void)devFree(dInfo);
  LSW_REQUIRE(mInfo == 0, LSW_ERR_NUMERIC, "factorization failed: zero or non-finite pivot");
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
  Geo2 g;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      g.intDom[a][b] = c.intDom[a][b];
      g.compDom[a][b] = c.compDom[a][b];
    }
  LSW_REQUIRE(gridInsidePml(g, c.x0, c.y0, c.hx, c.hy, c.nx, c.ny), LSW_ERR_ARG, "PML: Evaluation point out of range!");
  CellDev* dC = nullptr;
  double* dM = nullptr;
  double2* dCoef = nullptr;
  LSW_HIP(devMalloc(&dC, sizeof(CellDev)));
  LSW_HIP(devMalloc(&dM, N * sizeof(double)));
  LSW_HIP(devMalloc(&dCoef, 5 * N * sizeof(double2)));
  LSW_HIP(uploadSync(dC, &cd, sizeof(CellDev)));
  LSW_HIP(uploadSync(dM, m.data(), N * sizeof(double)));
  hipLaunchKernelGGL(k_assemble, dim3((unsigned)std::min<long long>((N + 255) / 256, 8192), 1), dim3(256), 0, s,
                     dC, dM, dCoef);
  LSW_HIP(hipGetLastError());
  // the temporaries are free for reuse once the assembly is done (a wait for
  // this stream, which set-up has just filled, instead of a later wait for
  // the whole device by the allocator)
  LSW_HIP(hipStreamSynchronize(s));
  IdleFrees idle;
  (void)devFree(dC);
  (void)devFree(dM);
  return dCoef;
}
}  // namespace lsw
