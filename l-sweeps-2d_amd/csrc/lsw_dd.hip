// DDCheckerboard on one GPU: geometry (src/DDCheckerboard.cpp:87-645), the
// L-sweeps preconditioner (:687-2608) and the global operator (:2611-2635).
#include "lsw_dd.h"

#include <chrono>
#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lsw_cplx.h"
#include "lsw_solver.h"

namespace lsw {

double2* assembleStencilDev(const CellDev& c, const std::vector<double>& m, hipStream_t s);

// --------------------------------------------------------------- constants
enum Ctx { kS1 = 0, kR = 1, kL = 2, kU = 3, kD = 4, kBL = 5, kTR = 6, kBR = 7, kTL = 8, kNumCtx = 9 };

// ------------------------------------------------------------------ kernels

// The local vectors of a step (right-hand sides X, solutions Xs) are laid
// out like the solve plan's tasks: slot s of the step's block lives at task
// t, lane q (tq[s] = t*W + q), and its point p at X[slotBase + p*W].
__device__ __forceinline__ long long slotBase(const int* __restrict__ tq, int sidx, int W, long long xs) {
  const int v = tq[sidx];
  const int t = v / W;
  return ((long long)t * xs) * W + (v - t * W);
}

// Right-hand side of sweep slot `sl` at point (i, j) of its cell (0 off the
// two trace sets): res0 = A01 v1 on set 0, res1 = -A10 v0 on set 1.
template <typename T2>
__device__ T2 sweepRhs(const SlotDev& sl, const CellDev& cd, int i, int j, const double2* __restrict__ coef,
                       const T2* __restrict__ T) {
  const int nx = cd.nx, ny = cd.ny, N = nx * ny;
  auto inSet = [&](int s, int ii, int jj) -> bool {
    return (ii == sl.setCol[s] && jj >= sl.setColLo[s] && jj < sl.setColHi[s]) ||
           (jj == sl.setRow[s] && ii >= sl.setRowLo[s] && ii < sl.setRowHi[s]);
  };
  auto val = [&](int s, int ii, int jj) -> T2 {
    if (sl.rowVal[s] >= 0 && jj == sl.setRow[s] && ii >= sl.setRowLo[s] && ii < sl.setRowHi[s])
      return T[sl.rowVal[s] + ii];
    return T[sl.colVal[s] + jj];
  };
  const bool in0 = inSet(0, i, j);
  const bool in1 = !in0 && inSet(1, i, j);
  if (!in0 && !in1) return CTraits<T2>::make(0, 0);
  const int other = in0 ? 1 : 0;
  const int ni[4] = {i - 1, i + 1, i, i};
  const int nj[4] = {j, j, j - 1, j + 1};
  const long long p = (long long)i * ny + j;
  double2 acc = make_double2(0.0, 0.0);
  for (int q = 0; q < 4; ++q) {
    if (ni[q] < 0 || ni[q] >= nx || nj[q] < 0 || nj[q] >= ny) continue;
    if (!inSet(other, ni[q], nj[q])) continue;
    const double2 a = coef[cd.coefOff + (long long)(q + 1) * N + p];
    const T2 vv = val(other, ni[q], nj[q]);
    acc = cfma(a, make_double2((double)vv.x, (double)vv.y), acc);
  }
  if (in1) acc = cneg(acc);
  return CTraits<T2>::make(acc.x, acc.y);
}

// Point t of trace-set segment m of a sweep slot (m: column 0, row 0,
// column 1, row 1); false past the segment's end.
__device__ __forceinline__ bool segmentPoint(const SlotDev& sl, int m, int t, int& i, int& j) {
  const int s = m >> 1;
  if ((m & 1) == 0) {
    if (sl.setCol[s] < 0 || t >= sl.setColHi[s] - sl.setColLo[s]) return false;
    i = sl.setCol[s];
    j = sl.setColLo[s] + t;
  } else {
    if (sl.setRow[s] < 0 || t >= sl.setRowHi[s] - sl.setRowLo[s]) return false;
    i = sl.setRowLo[s] + t;
    j = sl.setRow[s];
  }
  return true;
}

// Right-hand side of one local solve at point (i, j) = p of its cell.
//   stage 1: the windowed input restricted to the cell's owned set of its DD
//            (getSeparatedRHS + reduce/expand, src/DDCheckerboard.cpp:3564-3674)
//   sweeps : equivalent sources res0 = A01 v1 on set 0, res1 = -A10 v0 on set 1
//            (SubdomainStructured::getEquivalentSources/extractLTraces,
//            src/SubdomainStructured.cpp:1304-1539)
// Slot s of a block of right-hand sides is (slots[s % nSlots1], rhs = s / nSlots1);
// its trace storage is T + rhs*traceStride and its input vector xin + rhs*Nglob;
// a final solve also adds the summed sweep right-hand sides f.
template <typename T2>
__device__ T2 slotRhs(const SlotDev& sl, const CellDev& cd, int i, int j, const double2* __restrict__ xin,
                      const unsigned char* __restrict__ colT, const unsigned char* __restrict__ rowT, int Gy,
                      const double2* __restrict__ coef, const T2* __restrict__ T) {
  T2 out = CTraits<T2>::make(0, 0);
  if (sl.kind == 0) {
    const int slr = sl.d & 1, sud = sl.d >> 1;
    const bool own = i >= cd.vlr[slr][2] && i < cd.vlr[slr][5] && j >= cd.vud[sud][2] && j < cd.vud[sud][5];
    if (own) {
      const int gi = i + cd.ox, gj = j + cd.oy;
      const bool ct = colT[gi] != 0, rt = rowT[gj] != 0;
      const bool keep = (sl.d == 0) ? (!ct && !rt) : (sl.d == 1) ? (ct && !rt) : (sl.d == 2) ? (!ct && rt) : (ct && rt);
      if (keep) {
        const double2 v = xin[(long long)gi * Gy + gj];
        out = CTraits<T2>::make(v.x, v.y);
      }
    }
  } else {
    out = sweepRhs<T2>(sl, cd, i, j, coef, T);
  }
  return out;
}

// Stage-1 right-hand sides, one workgroup row per solve task: a thread takes
// one point and writes its W lanes as one contiguous vector (lanes past the
// task's right-hand sides get zeros); the lane loop is uniform, so a lane's
// slot description is one broadcast.  Each value also starts the final
// solve's right-hand side of its (d, cell) in F, which is laid out by the
// final step's plan (slot map tqF, width WF): the sweep steps add into it.
// When stage 1's plan has that same layout, F alone is written (writeX = 0)
// and the stage-1 solves read it.
template <typename T2, int W>
__global__ void k_build_rhs_tasks(const CellDev* __restrict__ cells, const SlotDev* __restrict__ slots, int nSlots1,
                                  const SolveTaskDev* __restrict__ tasks, long long traceStride, long long Nglob,
                                  const double2* __restrict__ xin, const unsigned char* __restrict__ colT,
                                  const unsigned char* __restrict__ rowT, int Gy, const double2* __restrict__ coef,
                                  const T2* __restrict__ T, T2* __restrict__ X, long long xs,
                                  T2* __restrict__ F, const int* __restrict__ tqF, int WF, int nSlotsF,
                                  int writeX) {
  const int t = blockIdx.y;
  const SolveTaskDev& tk = tasks[t];
  const int s0 = tk.slot[0] % nSlots1;
  const SlotDev& sl0 = slots[s0];
  const CellDev& cd = cells[sl0.cell];
  const int ny = cd.ny, N = cd.nx * ny;
  T2* x = X + (long long)t * xs * W;
  // A stage-1 task carries one (d, cell) for several right-hand sides: one
  // window test per point serves every lane, and when the lanes' places in
  // F are one task row of F's layout, the point's lanes are one vector store
  // (all W of them when the task fills the row or the row is this task's own).
  bool one = sl0.kind == 0;
  long long fb[W];
#pragma unroll
  for (int q = 0; q < W; ++q) {
    fb[q] = 0;
    if (q < tk.nrhs) {
      const int sidx = tk.slot[q];
      one = one && sidx % nSlots1 == s0;
      fb[q] = slotBase(tqF, (sidx / nSlots1) * nSlotsF + slots[sidx % nSlots1].grp, WF, xs);
    }
  }
  bool fvec = one && WF == W && (tk.nrhs == W || writeX == 0) && fb[0] % W == 0;
#pragma unroll
  for (int q = 1; q < W; ++q)
    if (q < tk.nrhs) fvec = fvec && fb[q] == fb[0] + q;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
    const int i = p / ny, j = p - i * ny;
    T2 out[W];
    if (one) {
      const int slr = sl0.d & 1, sud = sl0.d >> 1;
      const bool own = i >= cd.vlr[slr][2] && i < cd.vlr[slr][5] && j >= cd.vud[sud][2] && j < cd.vud[sud][5];
      const int gi = i + cd.ox, gj = j + cd.oy;
      bool keep = false;
      if (own) {
        const bool ct = colT[gi] != 0, rt = rowT[gj] != 0;
        keep = (sl0.d == 0) ? (!ct && !rt) : (sl0.d == 1) ? (ct && !rt) : (sl0.d == 2) ? (!ct && rt) : (ct && rt);
      }
#pragma unroll
      for (int q = 0; q < W; ++q) {
        out[q] = CTraits<T2>::make(0, 0);
        if (keep && q < tk.nrhs) {
          const double2 v = xin[(long long)(tk.slot[q] / nSlots1) * Nglob + (long long)gi * Gy + gj];
          out[q] = CTraits<T2>::make(v.x, v.y);
        }
      }
      if (fvec) {
        T2* f = F + fb[0] + (long long)p * WF;
#pragma unroll
        for (int q = 0; q < W; ++q) f[q] = out[q];
      } else {
#pragma unroll
        for (int q = 0; q < W; ++q)
          if (q < tk.nrhs) F[fb[q] + (long long)p * WF] = out[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < W; ++q) {
        out[q] = CTraits<T2>::make(0, 0);
        if (q < tk.nrhs) {
          const int sidx = tk.slot[q];
          const SlotDev& sl = slots[sidx % nSlots1];
          const int rhs = sidx / nSlots1;
          out[q] = slotRhs<T2>(sl, cd, i, j, xin + (long long)rhs * Nglob, colT, rowT, Gy, coef,
                               T + (long long)rhs * traceStride);
          F[fb[q] + (long long)p * WF] = out[q];
        }
      }
    }
    if (writeX) {
#pragma unroll
      for (int q = 0; q < W; ++q) x[(long long)p * W + q] = out[q];
    }
  }
}

// Sweep right-hand sides of the steps after the first, and their sum per
// (d, cell) into F (the final step's layout: slot map tqF, width WF, the
// final slot of a group being its index), in one pass: one workgroup per
// (group, right-hand side)
// writes the trace-set points of each of the group's slots (the buffers are
// zero elsewhere: cleared before the first sweep step, and each sweep step's
// k_post clears its own sets) and adds each point once into F, the slots one
// after the other.  By linearity the solutions of all solves of one (d, cell),
// which the reference adds into its result (`res +=`), are the solution of
// the sum of their right-hand sides, computed once per application.
template <typename T2>
__global__ void k_build_rhs_groups(const GroupDev* __restrict__ groups, int nGroups1,
                                   const CellDev* __restrict__ cells, const SlotDev* __restrict__ slots, int nSlots1,
                                   const int* __restrict__ tq, int W, long long traceStride,
                                   const double2* __restrict__ coef, const T2* __restrict__ T, T2* __restrict__ X,
                                   long long xs, T2* __restrict__ F, const int* __restrict__ tqF, int WF,
                                   int nSlotsF) {
  const GroupDev& g = groups[blockIdx.x % nGroups1];
  const int rhs = blockIdx.x / nGroups1;
  const T2* Tr = T + (long long)rhs * traceStride;
  for (int k = 0; k < g.n; ++k) {
    const SlotDev& sl = slots[g.slot[k]];
    if (sl.kind != 0) {
      const CellDev& cd = cells[sl.cell];
      const int ny = cd.ny;
      T2* x = X + slotBase(tq, rhs * nSlots1 + g.slot[k], W, xs);
      T2* f = F + slotBase(tqF, rhs * nSlotsF + sl.grp, WF, xs);
      auto onSeg = [&](int m, int i, int j) -> bool {
        const int s = m >> 1;
        if ((m & 1) == 0) return sl.setCol[s] >= 0 && i == sl.setCol[s] && j >= sl.setColLo[s] && j < sl.setColHi[s];
        return sl.setRow[s] >= 0 && j == sl.setRow[s] && i >= sl.setRowLo[s] && i < sl.setRowHi[s];
      };
      // the four segments' points as one index space over the workgroup (a
      // point on several segments gets the same value; only its first
      // segment adds it into F)
      int len[4], tot = 0;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int s = m >> 1;
        const bool col = (m & 1) == 0;
        const bool has = col ? sl.setCol[s] >= 0 : sl.setRow[s] >= 0;
        const int lo = col ? sl.setColLo[s] : sl.setRowLo[s], hi = col ? sl.setColHi[s] : sl.setRowHi[s];
        len[m] = has && hi > lo ? hi - lo : 0;
        tot += len[m];
      }
      for (int u = threadIdx.x; u < tot; u += blockDim.x) {
        int m = 0, t = u;
        while (t >= len[m]) t -= len[m++];
        const int s = m >> 1;
        const bool col = (m & 1) == 0;
        t += col ? sl.setColLo[s] : sl.setRowLo[s];
        const int i = col ? sl.setCol[s] : t, j = col ? t : sl.setRow[s];
        const T2 v = sweepRhs<T2>(sl, cd, i, j, coef, Tr);
        const long long p = (long long)i * ny + j;
        x[p * W] = v;
        bool dup = false;
        for (int mm = 0; mm < m; ++mm) dup = dup || onSeg(mm, i, j);
        if (!dup) f[p * WF] = cadd(f[p * WF], v);
      }
    }
    __syncthreads();
  }
}

// Trace extraction after the local solves of one step (assign or accumulate
// into the slot's own trace context), with the stage-3 overwrite of the
// incoming outer traces (src/DDCheckerboard.cpp:1861-1864,2085-2088,
// 2305-2308,2520-2526).
// Xs holds the solutions; the slot's trace sets in the right-hand-side
// buffer X are cleared (the next sweep step's sparse build relies on it).
template <typename T2>
__global__ void k_post(const CellDev* __restrict__ cells, const SlotDev* __restrict__ slots, int nSlots1,
                       const int* __restrict__ tq, int W, long long traceStride, const T2* __restrict__ Xs,
                       T2* __restrict__ X, long long xs, T2* __restrict__ T) {
  const int sidx = blockIdx.y;
  const SlotDev& sl = slots[sidx % nSlots1];
  T += (long long)(sidx / nSlots1) * traceStride;
  const CellDev& cd = cells[sl.cell];
  const int nx = cd.nx, ny = cd.ny;
  const long long base = slotBase(tq, sidx, W, xs);
  const T2* x = Xs + base;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl.kind != 0) {
    T2* xr = X + base;
    for (int m = 0; m < 4; ++m) {
      int i, j;
      if (segmentPoint(sl, m, t, i, j)) xr[((long long)i * ny + j) * W] = CTraits<T2>::make(0, 0);
    }
  }
  for (int o = 0; o < sl.nOut; ++o) {
    const bool isRow = sl.outIsRow[o] != 0;
    const int len = isRow ? nx : ny;
    if (t >= len) continue;
    const int i = isRow ? t : sl.outIdx[o];
    const int j = isRow ? sl.outIdx[o] : t;
    T2 v;
    if (j == sl.owRow)
      v = T[sl.owRowVal + i];
    else if (i == sl.owCol)
      v = T[sl.owColVal + j];
    else
      v = x[((long long)i * ny + j) * W];
    T2* d = T + sl.outDst[o] + t;
    *d = sl.outAdd[o] ? cadd(*d, v) : v;
  }
}

// Add the owned part of the final solutions to the DDs' global accumulators,
// one workgroup row per solve task: a thread takes one point and reads its W
// lanes as one contiguous vector.  Each lane is the final solve of one (DD,
// cell, right-hand side), so no two lanes or tasks add into the same
// accumulator entry; when the owned windows cover each point once (assign),
// the accumulators are written rather than added to (and not cleared first).
template <typename T2, int W>
__global__ void k_accumulate_tasks(const CellDev* __restrict__ cells, const SlotDev* __restrict__ slots, int nSlots1,
                                   const SolveTaskDev* __restrict__ tasks, const T2* __restrict__ Xs, long long xs,
                                   double2* __restrict__ Y, long long N, int Gy, int assign) {
  const int t = blockIdx.y;
  const SolveTaskDev& tk = tasks[t];
  const CellDev& cd = cells[slots[tk.slot[0] % nSlots1].cell];
  const int ny = cd.ny, Nc = cd.nx * ny;
  const T2* x = Xs + (long long)t * xs * W;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < Nc; p += gridDim.x * blockDim.x) {
    const int i = p / ny, j = p - i * ny;
    T2 v[W];
#pragma unroll
    for (int q = 0; q < W; ++q) v[q] = x[(long long)p * W + q];
#pragma unroll
    for (int q = 0; q < W; ++q) {
      if (q >= tk.nrhs) continue;
      const int sidx = tk.slot[q];
      const SlotDev& sl = slots[sidx % nSlots1];
      const int rhs = sidx / nSlots1;
      const int slr = sl.d & 1, sud = sl.d >> 1;
      if (i < cd.vlr[slr][2] || i >= cd.vlr[slr][5] || j < cd.vud[sud][2] || j >= cd.vud[sud][5]) continue;
      double2* y = Y + ((long long)rhs * 4 + sl.d) * N + (long long)(i + cd.ox) * Gy + (j + cd.oy);
      if (assign) {
        *y = make_double2((double)v[q].x, (double)v[q].y);
      } else {
        y->x += (double)v[q].x;
        y->y += (double)v[q].y;
      }
    }
  }
}

template <typename T2>
__global__ void k_copy_traces(const CopyOp* __restrict__ ops, int nOps1, long long traceStride,
                              T2* __restrict__ T, int tmax) {
  const CopyOp op = ops[blockIdx.y % nOps1];
  T += (long long)(blockIdx.y / nOps1) * traceStride;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tmax; t += gridDim.x * blockDim.x)
    T[op.dst + t] = T[op.src + t];
}

// res = res1 + (res2 + res3 + res4) (src/DDCheckerboard.cpp:733-735), per right-hand side
// zero n 8-byte words (a kernel rather than a memset node, so that the
// sequence can be captured into a graph)
__global__ void k_zero_words(unsigned long long* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

__global__ void k_combine(const double2* __restrict__ Y, double2* __restrict__ out, long long N, int nrhs) {
  const long long tot = N * nrhs;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (long long)gridDim.x * blockDim.x) {
    const long long r = q / N, g = q - r * N;
    const double2* y = Y + r * 4 * N;
    const double2 a = y[g], b = y[N + g], c = y[2 * N + g], d = y[3 * N + g];
    const double2 s = cadd(cadd(b, c), d);
    out[q] = cadd(a, s);
  }
}

// The final accumulation and combination in one pass, used when the cells'
// owned windows partition the grid for every DD: each global point finds,
// per DD d, the cell that owns it and that cell's final solve in the last
// step (slot d * nc + cell), converts the four values to FP64 and adds them
// as res1 + (res2 + res3 + res4) (src/DDCheckerboard.cpp:733-735).  Owner
// tables: colJ / colLi [2][Gx], rowI / rowLj [2][Gy], cellAt [nR][nC].
template <typename T2, int W>
__global__ __launch_bounds__(256) void k_gather_final(const CellDev* __restrict__ cells, const int* __restrict__ own,
                               const int* __restrict__ tq, long long xs, int nSlots1, int nc, int nC,
                               const T2* __restrict__ Xs, double2* __restrict__ out, int Gx, int Gy, int nrhs) {
  const long long N = (long long)Gx * Gy;
  const int* colJ = own;
  const int* colLi = own + 2 * Gx;
  const int* rowI = own + 4 * Gx;
  const int* rowLj = own + 4 * Gx + 2 * Gy;
  const int* cellAt = own + 4 * Gx + 4 * Gy;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < N;
       g += (long long)gridDim.x * blockDim.x) {
    const int gi = (int)(g / Gy), gj = (int)(g - (long long)gi * Gy);
    int sl[4];
    long long pw[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int slr = d & 1, sud = d >> 1;
      const int c = cellAt[rowI[sud * Gy + gj] * nC + colJ[slr * Gx + gi]];
      sl[d] = d * nc + c;
      pw[d] = ((long long)colLi[slr * Gx + gi] * cells[c].ny + rowLj[sud * Gy + gj]) * W;
    }
    // a task's lanes of one point are one contiguous W-vector: each DD keeps
    // the vector it loaded last, and every (right-hand side, DD) value it
    // holds is taken from registers (the right-hand sides of one (d, cell)
    // fill a task in turn, so each vector is loaded once)
    T2 lanes[4][W];
    long long at[4] = {-1, -1, -1, -1};
    for (int r = 0; r < nrhs; ++r) {
      double2 v[4];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const int tqv = tq[r * nSlots1 + sl[d]];
        const int t = tqv / W, q = tqv - t * W;
        const long long base = (long long)t * xs * W + pw[d];
        if (base != at[d]) {
#pragma unroll
          for (int k = 0; k < W; ++k) lanes[d][k] = Xs[base + k];
          at[d] = base;
        }
        T2 a = lanes[d][0];
#pragma unroll
        for (int k = 1; k < W; ++k)
          if (q == k) a = lanes[d][k];
        v[d] = make_double2((double)a.x, (double)a.y);
      }
      out[r * N + g] = cadd(v[0], cadd(cadd(v[1], v[2]), v[3]));
    }
  }
}

// Global 5-point operator y_r = A x_r for nrhs vectors stored one after the
// other (stride N): a point's five coefficients are read once and applied to
// every vector.
__global__ void k_matvec(const double2* __restrict__ coef, const double2* __restrict__ x,
                         double2* __restrict__ y, int Gx, int Gy, int nrhs) {
  const long long N = (long long)Gx * Gy;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < N;
       g += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(g / Gy), j = (int)(g - (long long)i * Gy);
    const double2 c0 = coef[g];
    const double2 zero = make_double2(0.0, 0.0);
    const double2 cw = i > 0 ? coef[N + g] : zero, ce = i < Gx - 1 ? coef[2 * N + g] : zero;
    const double2 cs = j > 0 ? coef[3 * N + g] : zero, cn = j < Gy - 1 ? coef[4 * N + g] : zero;
    for (int r = 0; r < nrhs; ++r) {
      const double2* xr = x + r * N;
      double2 acc = cmul(c0, xr[g]);
      if (i > 0) acc = cfma(cw, xr[g - Gy], acc);
      if (i < Gx - 1) acc = cfma(ce, xr[g + Gy], acc);
      if (j > 0) acc = cfma(cs, xr[g - 1], acc);
      if (j < Gy - 1) acc = cfma(cn, xr[g + 1], acc);
      y[r * N + g] = acc;
    }
  }
}

// Point-source RHS on the global grid (RHSSrcPts::eval, src/RHSSrcPts.cpp:87-112).
__global__ void k_point_sources(const double* __restrict__ src, int nsrc, double eps, double x0, double hx,
                                double y0, double hy, int Gx, int Gy, double2* __restrict__ b) {
  const long long N = (long long)Gx * Gy;
  const double pi = 3.14159265358979323846;
  const double scale = 1.0 / (pi * (eps * eps));
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < N;
       g += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(g / Gy), j = (int)(g - (long long)i * Gy);
    const double x = x0 + (double)(i + 1) * hx, y = y0 + (double)(j + 1) * hy;
    double acc = 0.0;
    for (int s = 0; s < nsrc; ++s) {
      const double dx = src[2 * s] - x, dy = src[2 * s + 1] - y;
      acc += scale * exp(-(dx * dx + dy * dy) / (eps * eps));
    }
    b[g] = make_double2(acc, 0.0);
  }
}

// ------------------------------------------------------------------ setup
DDCheckerboard::DDCheckerboard() = default;

DDCheckerboard::~DDCheckerboard() {
  // drain every stream that used the DD's memory (a caller's own stream: the
  // whole device), so that its blocks can be reused at once
  if (mForeignStream)
    (void)hipDeviceSynchronize();
  else if (mStream)
    (void)hipStreamSynchronize(mStream);
  IdleFrees idle;
  for (auto& kv : mGraphs) (void)hipGraphExecDestroy(kv.second);
  mGraphs.clear();
  if (mOwnEvent) (void)hipEventDestroy(mOwnEvent);
  for (void* p : {dX, dXs, dR, dG, dT, dF}) if (p) (void)devFree(p);
  if (dY) (void)devFree(dY);
  if (dGlobalCoef) (void)devFree(dGlobalCoef);
  mSteps.clear();  // their solve plans
  mArena.release();
  mFactors.reset();
  if (mStream) (void)hipStreamDestroy(mStream);
}

namespace {
// assignIndexSets: first limit the coordinate lies below, else 10
// (src/SubdomainStructured.cpp:1019-1116).  Classes are monotone along the
// axis, so each raw set is a range; s[k] = number of DOFs with class < k.
void rawStarts(const std::vector<double>& coords, const double lims[10], int s[12]) {
  std::vector<int> cls(coords.size());
  for (size_t i = 0; i < coords.size(); ++i) {
    int c = 10;
    for (int k = 0; k < 10; ++k)
      if (coords[i] < lims[k]) {
        c = k;
        break;
      }
    cls[i] = c;
  }
  for (size_t i = 1; i < cls.size(); ++i)
    LSW_REQUIRE(cls[i] >= cls[i - 1], LSW_ERR_ARG, "index sets are not monotone along the axis");
  for (int k = 0; k <= 11; ++k) {
    int n = 0;
    for (int c : cls) n += (c < k);
    s[k] = n;
  }
}
// 7-set views (src/SubdomainStructured.cpp:749-801).
void sevenView(const int s[12], int n, int v[2][8]) {
  const int un[8] = {0, s[1], s[2], s[3], s[6], s[7], s[8], n};
  const int sh[8] = {0, s[3], s[4], s[5], s[8], s[9], s[10], n};
  for (int k = 0; k < 8; ++k) {
    v[0][k] = un[k];
    v[1][k] = sh[k];
  }
}
}  // namespace

void DDCheckerboard::setUp(const Assembler* inputAssembler, const LocalLinearSolver* solver,
                           std::vector<DoubleVector> domain, std::vector<UInt> nCells,
                           std::vector<UInt> nElsInCell) {
  const auto t0 = std::chrono::steady_clock::now();
  double up0, al0;
  long long nu0, na0;
  threadMemTimes(&up0, &nu0, &al0, &na0);
  setUpGeometry(inputAssembler, solver, domain, nCells, nElsInCell);
  mSetupLog.emplace_back("geometry", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  initDevice();
  if (std::getenv("LSW_SETUP_LOG")) {
    double up1, al1;
    long long nu1, na1;
    threadMemTimes(&up1, &nu1, &al1, &na1);
    for (const auto& kv : mSetupLog) std::fprintf(stderr, "setup %-10s %.4f s\n", kv.first.c_str(), kv.second);
    std::fprintf(stderr, "setup uploads %lld %.4f s allocs %lld %.4f s\n", nu1 - nu0, up1 - up0, na1 - na0, al1 - al0);
  }
}

void DDCheckerboard::setUpDevice() {
  LSW_REQUIRE(!mCells.empty(), LSW_ERR_STATE, "setUpDevice before setUpGeometry");
  LSW_REQUIRE(mFactors == nullptr, LSW_ERR_STATE, "setUpDevice: already set up");
  initDevice();
}

void DDCheckerboard::setUpGeometry(const Assembler* inputAssembler, const LocalLinearSolver* solver,
                                   std::vector<DoubleVector> domain, std::vector<UInt> nCells,
                                   std::vector<UInt> nElsInCell) {
  const auto* a0 = dynamic_cast<const AssemblerHelmholtz2DFD*>(inputAssembler);
  LSW_REQUIRE(a0 != nullptr, LSW_ERR_ARG, "DDCheckerboard needs an AssemblerHelmholtz2DFD");
  LSW_REQUIRE(domain.size() == 2 && nCells.size() == 2 && nElsInCell.size() == 2, LSW_ERR_ARG,
              "DDCheckerboard: dimension must be 2");
  const auto* hs = dynamic_cast<const LocalLinearSolverHIP*>(solver);
  LSW_REQUIRE(hs != nullptr, LSW_ERR_ARG, "DDCheckerboard (HIP) needs a LocalLinearSolverHIP");
  mPrecision = hs->precision();
  const UInt order = a0->getOrder();
  LSW_REQUIRE(order == 2, LSW_ERR_UNSUPPORTED, "only Order 2 is supported");
  const UInt tt = (UInt)std::ceil(order / 2.0);
  LSW_REQUIRE(nElsInCell[0] > 4 * tt && nElsInCell[1] > 4 * tt, LSW_ERR_ARG, "Not enough elements in cell!");
  mDomain = domain;
  mNR = (int)nCells[0];
  mNC = (int)nCells[1];
  LSW_REQUIRE(mNR >= 1 && mNC >= 1, LSW_ERR_ARG, "DDCheckerboard: empty board");
  // Reference mode (src/DDCheckerboard.cpp:119-124): the x spacing comes from
  // the row count and NX.  It tiles consistently only when
  // NCellsX*(NX+1) == NCellsY*(NY+1); general mode takes the x spacing from the
  // column count and NY (the same grid whenever the reference's is consistent).
  const UInt n1 = mGeneral ? nCells[1] * nElsInCell[1] + nCells[1] - 1
                           : nCells[0] * nElsInCell[0] + nCells[0] - 1;
  const UInt n2 = mGeneral ? nCells[0] * nElsInCell[0] + nCells[0] - 1
                           : nCells[1] * nElsInCell[1] + nCells[1] - 1;
  LSW_REQUIRE(mGeneral || nCells[0] * (nElsInCell[0] + 1) == nCells[1] * (nElsInCell[1] + 1), LSW_ERR_UNSUPPORTED,
              "inconsistent tiling: NCellsX*(NX+1) != NCellsY*(NY+1) (use the general geometry mode)");
  const double h1 = (domain[0][1] - domain[0][0]) / n1;
  const double h2 = (domain[1][1] - domain[1][0]) / n2;
  const UInt NX = nElsInCell[0], NY = nElsInCell[1];
  mCells.clear();
  mCells.resize((size_t)mNR * mNC);
  for (int I = 0; I < mNR; ++I)
    for (int J = 0; J < mNC; ++J) {
      CellGeom& cg = mCells[(size_t)I * mNC + J];
      cg.I = I;
      cg.J = J;
      double lims1[10] = {}, lims2[10] = {};
      std::vector<DoubleVector> intDom(2, DoubleVector(2));
      double a = domain[0][0] + (NY + 1) * (UInt)J * h1;
      if (J > 0) {
        lims1[0] = a - tt * h1 - 0.5 * h1;
        lims1[1] = a - 0.5 * h1;
        lims1[2] = a + tt * h1 - 0.5 * h1;
        lims1[3] = a + 2 * tt * h1 - 0.5 * h1;
        lims1[4] = a + 3 * tt * h1 - 0.5 * h1;
        a -= 2 * tt * h1;
        if (a < domain[0][0]) a = domain[0][0];
      }
      double b = domain[0][0] + (NY + 1) * (UInt)J * h1 + NY * h1;
      if (J < mNC - 1) {
        lims1[5] = b - tt * h1 + 0.5 * h1;
        lims1[6] = b + 0.5 * h1;
        lims1[7] = b + tt * h1 + 0.5 * h1;
        lims1[8] = b + 2 * tt * h1 + 0.5 * h1;
        lims1[9] = b + 3 * tt * h1 + 0.5 * h1;
        b += 4 * tt * h1;
        if (b > domain[0][1]) b = domain[0][1];
      }
      intDom[0][0] = a;
      intDom[0][1] = b;
      a = domain[1][0] + (NX + 1) * (UInt)I * h2;
      if (I > 0) {
        lims2[0] = a - tt * h2 - 0.5 * h2;
        lims2[1] = a - 0.5 * h2;
        lims2[2] = a + tt * h2 - 0.5 * h2;
        lims2[3] = a + 2 * tt * h2 - 0.5 * h2;
        lims2[4] = a + 3 * tt * h2 - 0.5 * h2;
        a -= 2 * tt * h2;
        if (a < domain[1][0]) a = domain[1][0];
      }
      b = domain[1][0] + (NX + 1) * (UInt)I * h2 + NX * h2;
      if (I < mNR - 1) {
        lims2[5] = b - tt * h2 + 0.5 * h2;
        lims2[6] = b + 0.5 * h2;
        lims2[7] = b + tt * h2 + 0.5 * h2;
        lims2[8] = b + 2 * tt * h2 + 0.5 * h2;
        lims2[9] = b + 3 * tt * h2 + 0.5 * h2;
        b += 4 * tt * h2;
        if (b > domain[1][1]) b = domain[1][1];
      }
      intDom[1][0] = a;
      intDom[1][1] = b;
      std::vector<UInt> n(2);
      n[0] = (UInt)((intDom[0][1] - intDom[0][0] + 0.5 * h1) / h1);
      n[1] = (UInt)((intDom[1][1] - intDom[1][0] + 0.5 * h2) / h2);
      cg.asmb = *a0;
      cg.asmb.setDomain(intDom);
      cg.asmb.setN(n);
      cg.asmb.setOrder(order);
      cg.nx = cg.asmb.nxDofs();
      cg.ny = cg.asmb.nyDofs();
      // quirky min/max scan over the DOF points (src/DDCheckerboard.cpp:347-361)
      double minX = 0.0, maxX = 0.0, minY = 0.0, maxY = 0.0;
      for (int i = 0; i < cg.nx; ++i)
        for (int j = 0; j < cg.ny; ++j) {
          const double px = cg.asmb.xCoord(i), py = cg.asmb.yCoord(j);
          if (px < minX) minX = px;
          else if (px > maxX) maxX = px;
          if (py < minY) minY = py;
          else if (py > maxY) maxY = py;
        }
      if (J == 0)
        for (int k = 0; k < 5; ++k) lims1[k] = minX - h1;
      if (J == mNC - 1)
        for (int k = 5; k < 10; ++k) lims1[k] = maxX + h1;
      if (I == 0)
        for (int k = 0; k < 5; ++k) lims2[k] = minY - h2;
      if (I == mNR - 1)
        for (int k = 5; k < 10; ++k) lims2[k] = maxY + h2;
      std::memcpy(cg.lims[0], lims1, sizeof(lims1));
      std::memcpy(cg.lims[1], lims2, sizeof(lims2));
      std::vector<double> xs(cg.nx), ys(cg.ny);
      for (int i = 0; i < cg.nx; ++i) xs[i] = cg.asmb.xCoord(i);
      for (int j = 0; j < cg.ny; ++j) ys[j] = cg.asmb.yCoord(j);
      int sx[12], sy[12];
      rawStarts(xs, lims1, sx);
      rawStarts(ys, lims2, sy);
      sevenView(sx, cg.nx, cg.vlr);
      sevenView(sy, cg.ny, cg.vud);
      for (int f = 0; f < 2; ++f)
        for (int k = 1; k < 6; ++k) {
          if (k == 3) continue;
          LSW_REQUIRE(cg.vlr[f][k + 1] - cg.vlr[f][k] <= 1 && cg.vud[f][k + 1] - cg.vud[f][k] <= 1,
                      LSW_ERR_UNSUPPORTED, "trace sets thicker than one grid line");
        }
      cg.meshSize = cg.asmb.getMeshSize();
    }
  // global grid = the local grid of a 1x1 DD over the whole domain
  mGlobal = CellGeom();
  mGlobal.asmb = *a0;
  mGlobal.asmb.setDomain(domain);
  mGlobal.asmb.setN({n1, n2});
  mGlobal.asmb.setOrder(order);
  mGx = mGlobal.asmb.nxDofs();
  mGy = mGlobal.asmb.nyDofs();
  mGlobal.nx = mGx;
  mGlobal.ny = mGy;
  mMeshSize = mCells[0].meshSize;
  for (auto& c : mCells) mMeshSize = std::min(mMeshSize, c.meshSize);
  // local -> global offsets; owned DOFs must sit on global grid points
  std::vector<int> colOwner(mGx, -1), rowOwner(mGy, -1);
  std::vector<unsigned char> colT(mGx, 0), rowT(mGy, 0);
  for (auto& c : mCells) {
    c.ox = (int)std::lround((c.asmb.xCoord(0) - mGlobal.asmb.xCoord(0)) / h1);
    c.oy = (int)std::lround((c.asmb.yCoord(0) - mGlobal.asmb.yCoord(0)) / h2);
    for (int i = c.vlr[0][2]; i < c.vlr[0][5]; ++i) {
      const int gi = i + c.ox;
      LSW_REQUIRE(gi >= 0 && gi < mGx, LSW_ERR_ARG, "owned column outside the global grid");
      LSW_REQUIRE(std::fabs(c.asmb.xCoord(i) - mGlobal.asmb.xCoord(gi)) < 1e-6 * h1, LSW_ERR_ARG,
                  "local and global grids are not aligned");
      if (c.I == 0) {
        LSW_REQUIRE(colOwner[gi] < 0, LSW_ERR_ARG, "owned columns overlap");
        colOwner[gi] = c.J;
        colT[gi] = (i < c.vlr[0][3] || i >= c.vlr[0][4]) ? 1 : 0;
      }
    }
    for (int j = c.vud[0][2]; j < c.vud[0][5]; ++j) {
      const int gj = j + c.oy;
      LSW_REQUIRE(gj >= 0 && gj < mGy, LSW_ERR_ARG, "owned row outside the global grid");
      LSW_REQUIRE(std::fabs(c.asmb.yCoord(j) - mGlobal.asmb.yCoord(gj)) < 1e-6 * h2, LSW_ERR_ARG,
                  "local and global grids are not aligned");
      if (c.J == 0) {
        LSW_REQUIRE(rowOwner[gj] < 0, LSW_ERR_ARG, "owned rows overlap");
        rowOwner[gj] = c.I;
        rowT[gj] = (j < c.vud[0][3] || j >= c.vud[0][4]) ? 1 : 0;
      }
    }
  }
  for (int gi = 0; gi < mGx; ++gi) LSW_REQUIRE(colOwner[gi] >= 0, LSW_ERR_ARG, "owned columns do not cover the grid");
  for (int gj = 0; gj < mGy; ++gj) LSW_REQUIRE(rowOwner[gj] >= 0, LSW_ERR_ARG, "owned rows do not cover the grid");
  // For every DD the owned windows of the cells normally cover each global
  // point exactly once; then the final accumulation assigns instead of adding
  // into cleared accumulators.
  {
    std::vector<unsigned char> cnt((size_t)mGx * mGy);
    mOwnedPartition = true;
    for (int d = 0; d < 4 && mOwnedPartition; ++d) {
      std::fill(cnt.begin(), cnt.end(), 0);
      const int slr = d & 1, sud = d >> 1;
      for (const auto& c : mCells)
        for (int i = c.vlr[slr][2]; i < c.vlr[slr][5]; ++i)
          for (int j = c.vud[sud][2]; j < c.vud[sud][5]; ++j) {
            const long long gi = i + c.ox, gj = j + c.oy;
            if (gi < 0 || gi >= mGx || gj < 0 || gj >= mGy) {
              mOwnedPartition = false;
              continue;
            }
            unsigned char& k = cnt[(size_t)gi * mGy + gj];
            k = (unsigned char)std::min(2, k + 1);
          }
      for (unsigned char k : cnt) mOwnedPartition = mOwnedPartition && k == 1;
    }
  }
  // owner tables of the final gather: the partition must be a product of
  // per-column and per-row ownership (every cell of a board column owns the
  // same global columns at the same local indices, likewise for rows)
  // (LSW_STAGE1_IN_F=0 builds stage 1 in X; LSW_FINAL_GATHER=0 keeps the
  // accumulate-and-combine path)
  const char* s1Env = std::getenv("LSW_STAGE1_IN_F");
  mStage1InF = !(s1Env && std::atoi(s1Env) == 0);
  const char* gatherEnv = std::getenv("LSW_FINAL_GATHER");
  mGatherFinal = mOwnedPartition && mNR * mNC == (int)mCells.size() && !(gatherEnv && std::atoi(gatherEnv) == 0);
  mOwnHost.clear();
  if (mGatherFinal) {
    const size_t oLi = 2 * (size_t)mGx, oI = 4 * (size_t)mGx, oLj = oI + 2 * (size_t)mGy, oCell = oI + 4 * (size_t)mGy;
    mOwnHost.assign(oCell + (size_t)mNR * mNC, -1);
    auto claim = [&](size_t at, size_t atLocal, int owner, int local) {
      if (mOwnHost[at] < 0) {
        mOwnHost[at] = owner;
        mOwnHost[atLocal] = local;
      } else if (mOwnHost[at] != owner || mOwnHost[atLocal] != local) {
        mGatherFinal = false;
      }
    };
    for (size_t ci = 0; ci < mCells.size() && mGatherFinal; ++ci) {
      const auto& c = mCells[ci];
      if (c.I < 0 || c.I >= mNR || c.J < 0 || c.J >= mNC || mOwnHost[oCell + (size_t)c.I * mNC + c.J] >= 0) {
        mGatherFinal = false;
        break;
      }
      mOwnHost[oCell + (size_t)c.I * mNC + c.J] = (int)ci;
      for (int h = 0; h < 2; ++h) {
        for (int i = c.vlr[h][2]; i < c.vlr[h][5]; ++i)
          claim((size_t)h * mGx + (i + c.ox), oLi + (size_t)h * mGx + (i + c.ox), c.J, i);
        for (int j = c.vud[h][2]; j < c.vud[h][5]; ++j)
          claim(oI + (size_t)h * mGy + (j + c.oy), oLj + (size_t)h * mGy + (j + c.oy), c.I, j);
      }
    }
    for (int v : mOwnHost) mGatherFinal = mGatherFinal && v >= 0;
    if (!mGatherFinal) mOwnHost.clear();
  }
  // neighbours share the perpendicular grid
  for (auto& c : mCells) {
    if (c.J + 1 < mNC) {
      const auto& r = mCells[(size_t)c.I * mNC + c.J + 1];
      LSW_REQUIRE(r.ny == c.ny && r.oy == c.oy, LSW_ERR_ARG, "row neighbours have different y grids");
    }
    if (c.I + 1 < mNR) {
      const auto& t = mCells[(size_t)(c.I + 1) * mNC + c.J];
      LSW_REQUIRE(t.nx == c.nx && t.ox == c.ox, LSW_ERR_ARG, "column neighbours have different x grids");
    }
  }
  mColT = colT;
  mRowT = rowT;
}

void DDCheckerboard::initDevice() {
  // non-blocking: the synchronous copies of other problems' set-up (on the
  // legacy stream, from other server workers) must not wait for this
  // problem's solves
  if (!mStream) LSW_HIP(hipStreamCreateWithFlags(&mStream, hipStreamNonBlocking));
  dColTrace = mArena.put(mColT);
  dRowTrace = mArena.put(mRowT);

  auto toDev = [&](const CellGeom& c) {
    CellDev d;
    std::memset(&d, 0, sizeof(d));
    d.nx = c.nx;
    d.ny = c.ny;
    d.ox = c.ox;
    d.oy = c.oy;
    std::memcpy(d.vlr, c.vlr, sizeof(d.vlr));
    std::memcpy(d.vud, c.vud, sizeof(d.vud));
    d.x0 = c.asmb.x0();
    d.y0 = c.asmb.y0();
    d.hx = c.asmb.hx();
    d.hy = c.asmb.hy();
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        d.intDom[a][b] = c.asmb.intDomain()[a][b];
        d.compDom[a][b] = c.asmb.compDomain()[a][b];
      }
    d.C = c.asmb.getPML()->getAbsorptionCoeff();
    d.omega = c.asmb.getOmega();
    return d;
  };
  auto lap = [&](const char* what, std::chrono::steady_clock::time_point& t) {
    const auto n = std::chrono::steady_clock::now();
    mSetupLog.emplace_back(what, std::chrono::duration<double>(n - t).count());
    t = n;
  };
  auto t = std::chrono::steady_clock::now();
  mFactors = std::make_unique<FactorSet>();
  mFactors->setSolvePrecision(mPrecision);
  for (auto& c : mCells) {
    std::vector<double> m;
    c.asmb.squaredSlowness(m);
    mFactors->addCell(toDev(c), m);
  }
  lap("media", t);
  mFactors->setup(mStream);
  lap("trees", t);
  {
    std::vector<double> m;
    mGlobal.asmb.squaredSlowness(m);
    mGlobalDev = toDev(mGlobal);
    mGlobalDev.mOff = 0;
    mGlobalDev.coefOff = 0;
    dGlobalCoef = assembleStencilDev(mGlobalDev, m, mStream);
  }
  lap("global", t);
  buildSchedule();
  lap("schedule", t);
  if (mGatherFinal) {
    // the gather addresses the last step's slot of (DD d, cell c) as d * nc + c
    const int nc = (int)mCells.size();
    const Step& fin = mSteps.back();
    bool ok = fin.deferred && fin.hSlots.size() == (size_t)4 * nc;
    for (size_t k = 0; ok && k < fin.hSlots.size(); ++k)
      ok = fin.hSlots[k].d == (int)k / nc && fin.hSlots[k].cell == (int)k % nc;
    mGatherFinal = ok;
    if (ok) dOwn = mArena.put(mOwnHost);
  }
  // every descriptor of the schedule, the trace classes and owner tables
  mArena.flush();
  allocBuffers();
  markOwnWork();  // the stencil assembly and the buffer clears
  lap("buffers", t);
}

// ------------------------------------------------------------ the schedule
void DDCheckerboard::buildSchedule() {
  const int nc = (int)mCells.size();
  mTmax = 0;
  for (auto& c : mCells) mTmax = std::max(mTmax, std::max(c.nx, c.ny));
  mTraceEntries = 4LL * nc * kNumCtx * 8 * mTmax;
  auto tOff = [&](int d, int c, int ctx, int k) -> long long {
    return ((((long long)d * nc + c) * kNumCtx + ctx) * 8 + k) * (long long)mTmax;
  };
  auto cidx = [&](int I, int J) { return I * mNC + J; };

  struct HostStep {
    std::vector<SlotDev> slots;
    std::vector<CopyOp> copies;
  };
  std::vector<HostStep> steps;

  auto baseSlot = [&](int d, int kind, int c) {
    SlotDev s;
    std::memset(&s, 0, sizeof(s));
    s.d = d;
    s.kind = kind;
    s.cell = c;
    for (int k = 0; k < 2; ++k) {
      s.setCol[k] = s.setRow[k] = -1;
      s.colVal[k] = s.rowVal[k] = -1;
    }
    s.owCol = s.owRow = -1;
    s.owColVal = s.owRowVal = 0;
    return s;
  };
  auto addOut = [&](SlotDev& s, bool isRow, int idx, int lo, int hi, long long dst, bool add) {
    if (hi <= lo) return;  // empty trace set
    LSW_REQUIRE(s.nOut < 8, LSW_ERR_ARG, "too many trace outputs");
    s.outIsRow[s.nOut] = isRow ? 1 : 0;
    s.outIdx[s.nOut] = idx;
    s.outAdd[s.nOut] = add ? 1 : 0;
    s.outDst[s.nOut] = dst;
    ++s.nOut;
  };
  // set the four traces of one kind (LR: columns, UD: rows) from v (7-view)
  auto outFour = [&](SlotDev& s, bool isRow, const int v[8], int d, int c, int ctx, bool add) {
    const int sets[4] = {1, 2, 4, 5};
    for (int k = 0; k < 4; ++k)
      addOut(s, isRow, v[sets[k]], v[sets[k]], v[sets[k] + 1], tOff(d, c, ctx, (isRow ? 4 : 0) + k), add);
  };

  // ---- stage 1
  {
    HostStep st;
    for (int d = 0; d < 4; ++d)
      for (int c = 0; c < nc; ++c) {
        const auto& cg = mCells[c];
        const int slr = d & 1, sud = d >> 1;
        SlotDev s = baseSlot(d, 0, c);
        outFour(s, false, cg.vlr[slr], d, c, kS1, false);
        outFour(s, true, cg.vud[sud], d, c, kS1, false);
        st.slots.push_back(s);
        // copies for stage 2: S1.LR -> R.LR, L.LR ; S1.UD -> U.UD, D.UD
        for (int k = 0; k < 4; ++k) {
          st.copies.push_back({tOff(d, c, kR, k), tOff(d, c, kS1, k)});
          st.copies.push_back({tOff(d, c, kL, k), tOff(d, c, kS1, k)});
          st.copies.push_back({tOff(d, c, kU, 4 + k), tOff(d, c, kS1, 4 + k)});
          st.copies.push_back({tOff(d, c, kD, 4 + k), tOff(d, c, kS1, 4 + k)});
        }
      }
    steps.push_back(st);
  }
  // Global steps: stage 1 at step 0, stage-2 step t at step t, and stage-3
  // anti-diagonal t at step t + 2.  An L-sweep on cell (I, J) reads the
  // stage-2 traces of that cell from the sweeps that pass it at stage-2 steps
  // 1+a and 1+b (a + b = t), so they are final after step t + 1; the traces
  // are copied into the L-sweep's context at the end of step t + 1 (the
  // reference copies them after the whole of stage 2; they no longer change
  // after their step, so the values are the same).  This runs the two stages
  // side by side: 2n-1 steps instead of 3n-3 on an n x n board.
  const int nStage2 = std::max(mNR, mNC) - 1;
  const int nStage3 = (mNR >= 2 && mNC >= 2) ? (mNR - 2) + (mNC - 2) + 1 : 0;
  steps.resize((size_t)1 + std::max(nStage2, nStage3 > 0 ? nStage3 + 1 : 0));
  // ---- stage 2
  for (int t = 1; t <= nStage2; ++t) {
    HostStep& st = steps[t];
    for (int d = 0; d < 4; ++d) {
      const int slr = d & 1, sud = d >> 1;
      // right sweep: column t (src/DDCheckerboard.cpp:1168-1350)
      if (t < mNC)
        for (int I = 0; I < mNR; ++I) {
          const int c = cidx(I, t), l = cidx(I, t - 1);
          const auto& v = mCells[c].vlr[slr];
          SlotDev s = baseSlot(d, 1, c);
          s.setCol[0] = v[1]; s.setColLo[0] = 0; s.setColHi[0] = mCells[c].ny;
          s.setCol[1] = v[2]; s.setColLo[1] = 0; s.setColHi[1] = mCells[c].ny;
          s.colVal[0] = tOff(d, l, kR, 2);
          s.colVal[1] = tOff(d, l, kR, 3);
          outFour(s, true, mCells[c].vud[sud], d, c, kR, false);
          addOut(s, false, v[4], v[4], v[5], tOff(d, c, kR, 2), true);
          addOut(s, false, v[5], v[5], v[6], tOff(d, c, kR, 3), true);
          st.slots.push_back(s);
        }
      // left sweep: column nC-1-t (src/DDCheckerboard.cpp:976-1164)
      if (t < mNC)
        for (int I = 0; I < mNR; ++I) {
          const int J = mNC - 1 - t;
          const int c = cidx(I, J), r = cidx(I, J + 1);
          const auto& v = mCells[c].vlr[slr];
          SlotDev s = baseSlot(d, 2, c);
          s.setCol[0] = v[5]; s.setColLo[0] = 0; s.setColHi[0] = mCells[c].ny;
          s.setCol[1] = v[4]; s.setColLo[1] = 0; s.setColHi[1] = mCells[c].ny;
          s.colVal[0] = tOff(d, r, kL, 1);
          s.colVal[1] = tOff(d, r, kL, 0);
          outFour(s, true, mCells[c].vud[sud], d, c, kL, false);
          addOut(s, false, v[1], v[1], v[2], tOff(d, c, kL, 0), true);
          addOut(s, false, v[2], v[2], v[3], tOff(d, c, kL, 1), true);
          st.slots.push_back(s);
        }
      // up sweep: row t (src/DDCheckerboard.cpp:1544-1724)
      if (t < mNR)
        for (int J = 0; J < mNC; ++J) {
          const int c = cidx(t, J), bo = cidx(t - 1, J);
          const auto& v = mCells[c].vud[sud];
          SlotDev s = baseSlot(d, 3, c);
          s.setRow[0] = v[1]; s.setRowLo[0] = 0; s.setRowHi[0] = mCells[c].nx;
          s.setRow[1] = v[2]; s.setRowLo[1] = 0; s.setRowHi[1] = mCells[c].nx;
          s.rowVal[0] = tOff(d, bo, kU, 4 + 2);
          s.rowVal[1] = tOff(d, bo, kU, 4 + 3);
          outFour(s, false, mCells[c].vlr[slr], d, c, kU, false);
          addOut(s, true, v[4], v[4], v[5], tOff(d, c, kU, 4 + 2), true);
          addOut(s, true, v[5], v[5], v[6], tOff(d, c, kU, 4 + 3), true);
          st.slots.push_back(s);
        }
      // down sweep: row nR-1-t (src/DDCheckerboard.cpp:1353-1540)
      if (t < mNR)
        for (int J = 0; J < mNC; ++J) {
          const int I = mNR - 1 - t;
          const int c = cidx(I, J), to = cidx(I + 1, J);
          const auto& v = mCells[c].vud[sud];
          SlotDev s = baseSlot(d, 4, c);
          s.setRow[0] = v[5]; s.setRowLo[0] = 0; s.setRowHi[0] = mCells[c].nx;
          s.setRow[1] = v[4]; s.setRowLo[1] = 0; s.setRowHi[1] = mCells[c].nx;
          s.rowVal[0] = tOff(d, to, kD, 4 + 1);
          s.rowVal[1] = tOff(d, to, kD, 4 + 0);
          outFour(s, false, mCells[c].vlr[slr], d, c, kD, false);
          addOut(s, true, v[1], v[1], v[2], tOff(d, c, kD, 4 + 0), true);
          addOut(s, true, v[2], v[2], v[3], tOff(d, c, kD, 4 + 1), true);
          st.slots.push_back(s);
        }
    }
  }
  if (nStage2 == 0) {
    // 1 x 1 board: no sweeps at all
  }
  // ---- stage 3
  for (int t = 0; t < nStage3; ++t) {
    HostStep& st = steps[t + 2];
    for (int d = 0; d < 4; ++d) {
      const int slr = d & 1, sud = d >> 1;
      for (int kind = 5; kind <= 8; ++kind) {
        const bool fromLeft = (kind == 5 || kind == 8);    // BL, TL
        const bool fromBottom = (kind == 5 || kind == 7);  // BL, BR
        const int ctx = kind;                               // kBL..kTL == 5..8
        for (int a = 0; a <= t; ++a) {
          const int b = t - a;
          const int I = fromBottom ? 1 + a : mNR - 2 - a;
          const int J = fromLeft ? 1 + b : mNC - 2 - b;
          if (I < 0 || I >= mNR || J < 0 || J >= mNC) continue;
          if (fromBottom ? I < 1 : I > mNR - 2) continue;
          if (fromLeft ? J < 1 : J > mNC - 2) continue;
          const int c = cidx(I, J);
          const auto& cg = mCells[c];
          const int* vl = cg.vlr[slr];
          const int* vu = cg.vud[sud];
          const int nb = fromLeft ? cidx(I, J - 1) : cidx(I, J + 1);
          const int nv = fromBottom ? cidx(I - 1, J) : cidx(I + 1, J);
          SlotDev s = baseSlot(d, kind, c);
          const int oc = fromLeft ? vl[1] : vl[5], ic = fromLeft ? vl[2] : vl[4];
          const int orr = fromBottom ? vu[1] : vu[5], ir = fromBottom ? vu[2] : vu[4];
          s.setCol[0] = oc;
          s.setColLo[0] = fromBottom ? vu[1] : 0;
          s.setColHi[0] = fromBottom ? cg.ny : vu[6];
          s.setRow[0] = orr;
          s.setRowLo[0] = fromLeft ? vl[1] : 0;
          s.setRowHi[0] = fromLeft ? cg.nx : vl[6];
          s.setCol[1] = ic;
          s.setColLo[1] = fromBottom ? vu[2] : 0;
          s.setColHi[1] = fromBottom ? cg.ny : vu[5];
          s.setRow[1] = ir;
          s.setRowLo[1] = fromLeft ? vl[2] : 0;
          s.setRowHi[1] = fromLeft ? cg.nx : vl[5];
          // incoming: from the left neighbour's traceN/NP (LR 2,3) or the
          // right neighbour's trace0/1 (LR 0,1); bottom's UD 2,3 or top's UD 0,1
          const long long colOuter = fromLeft ? tOff(d, nb, ctx, 2) : tOff(d, nb, ctx, 1);
          const long long colInner = fromLeft ? tOff(d, nb, ctx, 3) : tOff(d, nb, ctx, 0);
          const long long rowOuter = fromBottom ? tOff(d, nv, ctx, 4 + 2) : tOff(d, nv, ctx, 4 + 1);
          const long long rowInner = fromBottom ? tOff(d, nv, ctx, 4 + 3) : tOff(d, nv, ctx, 4 + 0);
          s.colVal[0] = colOuter;
          s.rowVal[0] = rowOuter;
          s.colVal[1] = colInner;
          s.rowVal[1] = rowInner;
          s.owCol = oc;
          s.owColVal = colOuter;
          s.owRow = orr;
          s.owRowVal = rowOuter;
          if (fromLeft) {
            addOut(s, false, vl[4], vl[4], vl[5], tOff(d, c, ctx, 2), true);
            addOut(s, false, vl[5], vl[5], vl[6], tOff(d, c, ctx, 3), true);
          } else {
            addOut(s, false, vl[1], vl[1], vl[2], tOff(d, c, ctx, 0), true);
            addOut(s, false, vl[2], vl[2], vl[3], tOff(d, c, ctx, 1), true);
          }
          if (fromBottom) {
            addOut(s, true, vu[4], vu[4], vu[5], tOff(d, c, ctx, 4 + 2), true);
            addOut(s, true, vu[5], vu[5], vu[6], tOff(d, c, ctx, 4 + 3), true);
          } else {
            addOut(s, true, vu[1], vu[1], vu[2], tOff(d, c, ctx, 4 + 0), true);
            addOut(s, true, vu[2], vu[2], vu[3], tOff(d, c, ctx, 4 + 1), true);
          }
          st.slots.push_back(s);
        }
      }
    }
  }
  // The L-sweeps take their traces by value (include/DDCheckerboard.h:209-239):
  // UD traces from the horizontal sweep that reaches the cell (R from the
  // left, L from the right), LR traces from the vertical one (U from below, D
  // from above) -- for every cell, also those the L-sweep does not solve,
  // whose copies its neighbours read.  A cell's copy is made at the end of
  // the step after which both sources are final: R holds column J from step
  // J on (column 0: stage 1), L column J from step nC-1-J (last column:
  // stage 1), U row I from step I, D row I from step nR-1-I.
  if (nStage3 > 0) {
    for (int d = 0; d < 4; ++d)
      for (int kind = 5; kind <= 8; ++kind) {
        const bool fromLeft = (kind == 5 || kind == 8);
        const bool fromBottom = (kind == 5 || kind == 7);
        const int horiz = fromLeft ? kR : kL, vert = fromBottom ? kU : kD;
        for (int I = 0; I < mNR; ++I)
          for (int J = 0; J < mNC; ++J) {
            const int c = cidx(I, J);
            const int hs = fromLeft ? J : (J <= mNC - 2 ? mNC - 1 - J : 0);
            const int vs = fromBottom ? I : (I <= mNR - 2 ? mNR - 1 - I : 0);
            auto& cp = steps[std::max(hs, vs)].copies;
            for (int k = 0; k < 4; ++k) {
              cp.push_back({tOff(d, c, kind, 4 + k), tOff(d, c, horiz, 4 + k)});
              cp.push_back({tOff(d, c, kind, k), tOff(d, c, vert, k)});
            }
          }
      }
  }

  // ---- the final solves: one per (d, cell), right-hand side = the windowed
  // RHS of stage 1 plus the summed sweep right-hand sides of that (d, cell)
  {
    HostStep st;
    for (int d = 0; d < 4; ++d)
      for (int c = 0; c < nc; ++c) {
        SlotDev s = baseSlot(d, 0, c);
        s.deferred = 1;
        st.slots.push_back(s);
      }
    steps.push_back(st);
  }
  for (auto& hst : steps)
    for (auto& sl : hst.slots) sl.grp = sl.d * nc + sl.cell;
  // stage 1 starts every (d, cell)'s final right-hand side in F, addressed
  // through the final step's slot of the same index
  for (const HostStep* hs : {&steps.front(), &steps.back()}) {
    LSW_REQUIRE(hs->slots.size() == (size_t)4 * nc, LSW_ERR_STATE, "stage 1 or final step misses a (d, cell)");
    for (size_t k = 0; k < hs->slots.size(); ++k)
      LSW_REQUIRE(hs->slots[k].grp == (int)k, LSW_ERR_STATE, "stage 1 or final step out of (d, cell) order");
  }

  // slot buffers, tasks (same cell -> one task of up to kMaxRhs), groups
  mMaxSlots = 0;
  mSolvesPerPrecond = 0;
  mSolveBytesPerPrecond = -1;
  const double S = (mPrecision == LSW_FP64) ? 16.0 : 8.0;
  for (auto& st : steps) mMaxSlots = std::max(mMaxSlots, (int)st.slots.size());
  mSteps.clear();
  for (size_t si = 0; si < steps.size(); ++si) {
    auto& hst = steps[si];
    Step st;
    st.deferred = si + 1 == steps.size();
    for (size_t k = 0; k < hst.slots.size(); ++k) hst.slots[k].slot = (int)k;
    st.hSlots = hst.slots;
    for (const auto& sl : hst.slots) {
      // a final solve's right-hand side also holds the summed sweep sources,
      // which live on trace lines outside the window: no forward pruning
      st.hSupport.push_back(sl.deferred ? std::vector<Rect>{} : slotSupport(sl));
      // a sweep solve is only read on its output trace lines
      std::vector<Rect> need;
      const CellGeom& cg = mCells[sl.cell];
      for (int o = 0; o < sl.nOut; ++o) {
        if (sl.outIsRow[o])
          need.push_back({0, cg.nx, sl.outIdx[o], sl.outIdx[o] + 1});
        else
          need.push_back({sl.outIdx[o], sl.outIdx[o] + 1, 0, cg.ny});
      }
      st.hNeed.push_back(need);
    }
    (void)S;
    mSolvesPerPrecond += (long long)hst.slots.size();
    // groups: (d, cell)
    std::vector<GroupDev> groups;
    for (auto& s : hst.slots) {
      bool found = false;
      for (auto& g : groups)
        if (g.d == s.d && g.cell == s.cell) {
          LSW_REQUIRE(g.n < 4, LSW_ERR_ARG, "too many solves of one cell in a step");
          g.slot[g.n++] = s.slot;
          found = true;
          break;
        }
      if (!found) {
        GroupDev g;
        std::memset(&g, 0, sizeof(g));
        g.d = s.d;
        g.cell = s.cell;
        g.n = 1;
        g.slot[0] = s.slot;
        groups.push_back(g);
      }
    }
    st.nSlots = (int)hst.slots.size();
    st.nGroups = (int)groups.size();
    st.nCopies = (int)hst.copies.size();
    LSW_REQUIRE(st.nSlots <= 65535 && st.nCopies <= 65535 * 8, LSW_ERR_CAPACITY, "board too large for one launch");
    // (uploaded by the arena's flush at the end of the device set-up; the
    // solve plans are built when a block width is first solved)
    if (hst.slots.empty()) hst.slots.emplace_back();
    if (groups.empty()) groups.emplace_back();
    st.dSlots = mArena.put(hst.slots);
    st.dGroups = mArena.put(groups);
    st.dCopies = mArena.put(hst.copies);
    mSteps.push_back(std::move(st));
  }
}

double DDCheckerboard::solveBytesPerPrecond() const {
  if (mSolveBytesPerPrecond < 0) {
    double b = 0;
    for (const auto& st : mSteps) {
      const SolvePlan& pl = planFor(st, 1);
      b += pl.fwdBytes + pl.bwdBytes;
    }
    mSolveBytesPerPrecond = b;
  }
  return mSolveBytesPerPrecond;
}

// Solve plan of one step for a block of nrhs right-hand sides.  Slots are
// ordered by (cell, DD, sweep, right-hand side) and cut into tasks of up to
// kMaxRhs: with one right-hand side a task carries the four DDs of a cell
// (one factor read for all of them); with several, a task carries one DD and
// sweep for several right-hand sides, so its support stays a few grid lines.
const SolvePlan& DDCheckerboard::planFor(const Step& st, int nrhs) const {
  auto it = st.plans.find(nrhs);
  if (it != st.plans.end()) return *it->second;
  const int n1 = (int)st.hSlots.size();
  std::vector<int> order((size_t)n1 * nrhs);
  for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
  auto key = [&](int s) {
    const SlotDev& a = st.hSlots[s % n1];
    return std::make_tuple(a.cell, a.d, a.kind, s / n1, s % n1);
  };
  std::sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
  // one right-hand side: the four DDs' solves of a cell share a task (4);
  // blocks: up to 8 per task where the LDS staging allows
  const int perTask = nrhs == 1 ? kTaskRhs4 : mFactors->maxTaskRhs();
  std::vector<SolveTask> tasks;
  std::vector<std::vector<Rect>> supports, needs;
  for (size_t k = 0; k < order.size();) {
    SolveTask tk;
    std::memset(&tk, 0, sizeof(tk));
    tk.cell = st.hSlots[order[k] % n1].cell;
    std::vector<Rect> sup, need;
    while (k < order.size() && st.hSlots[order[k] % n1].cell == tk.cell && tk.nrhs < perTask) {
      const int s = order[k++];
      tk.slot[tk.nrhs++] = s;
      const auto& r = st.hSupport[s % n1];
      sup.insert(sup.end(), r.begin(), r.end());
      const auto& nd = st.hNeed[s % n1];
      need.insert(need.end(), nd.begin(), nd.end());
    }
    // (the right-hand sides of one slot geometry repeat its rectangles)
    auto dedupe = [](std::vector<Rect>& v) {
      auto key = [](const Rect& r) { return std::make_tuple(r.i0, r.i1, r.j0, r.j1); };
      std::sort(v.begin(), v.end(), [&](const Rect& a, const Rect& b) { return key(a) < key(b); });
      v.erase(std::unique(v.begin(), v.end(), [&](const Rect& a, const Rect& b) { return key(a) == key(b); }), v.end());
    };
    dedupe(sup);
    dedupe(need);
    tasks.push_back(tk);
    supports.push_back(std::move(sup));
    needs.push_back(std::move(need));
  }
  auto plan = std::make_unique<SolvePlan>();
  if (std::getenv("LSW_PLAN_TIME"))
    std::fprintf(stderr, "planFor nrhs %d slots %d tasks %zu\n", nrhs, n1, tasks.size());
  // the final solves need the whole solution; the sweep solves only their traces
  mFactors->buildPlan(tasks, supports, *plan, st.deferred ? nullptr : &needs, /*outOfPlace=*/true,
                      &shapeFor(nrhs));
  const SolvePlan& ref = *plan;
  st.plans[nrhs] = std::move(plan);
  return ref;
}

int DDCheckerboard::builtWidth(int lo, int hi) const {
  if (mSteps.empty()) return 0;
  // (every step's plan for a width is built by the first application at it)
  const auto& plans = mSteps.back().plans;
  auto it = plans.lower_bound(lo);
  return it != plans.end() && it->first <= hi ? it->first : 0;
}

// Rows of the task-major solve scratch (R, G) a block of nrhs right-hand
// sides needs in its largest step: planFor cuts each cell's slots into tasks
// of perTask, and every task takes taskWidth(largest task) rows.
long long DDCheckerboard::taskRowsFor(int nrhs) const {
  const int perTask = nrhs == 1 ? kTaskRhs4 : mFactors->maxTaskRhs();
  long long best = 0;
  std::vector<long long> perCell(mCells.size());
  for (const auto& st : mSteps) {
    std::fill(perCell.begin(), perCell.end(), 0);
    for (const auto& sl : st.hSlots) perCell[sl.cell] += nrhs;
    long long tasks = 0, widest = 0;
    for (long long c : perCell) {
      tasks += (c + perTask - 1) / perTask;
      widest = std::max(widest, std::min<long long>(c, perTask));
    }
    best = std::max(best, tasks * taskWidth((int)std::max<long long>(widest, 1)));
  }
  return best;
}

void DDCheckerboard::ensureBuffers(int nrhs) const {
  if (nrhs <= mBufRhs) return;
  const size_t es = (mPrecision == LSW_FP64) ? sizeof(double2) : sizeof(float2);
  {
    // buffers used on the DD's own stream only: drain it, and the blocks are
    // reusable at once
    const bool own = !mForeignStream && mStream && (dX || dY);
    if (own) LSW_HIP(hipStreamSynchronize(mStream));
    std::unique_ptr<IdleFrees> idle(own ? new IdleFrees : nullptr);
    for (void* p : {dX, dXs, dR, dG, dT, dF}) if (p) (void)devFree(p);
    if (dY) (void)devFree(dY);
  }
  dX = dXs = dR = dG = dT = dF = nullptr;
  dY = nullptr;
  mBufRhs = 0;
  // X, Xs, R and G serve every block width up to nrhs
  long long rows = 0;
  for (int n = 1; n <= nrhs; ++n) rows = std::max(rows, taskRowsFor(n));
  mTaskRows = rows;
  LSW_HIP(devMalloc(&dF, (size_t)rows * mXs * es));
  LSW_HIP(devMalloc(&dX, (size_t)rows * mXs * es));
  LSW_HIP(devMalloc(&dXs, (size_t)rows * mXs * es));
  LSW_HIP(devMalloc(&dR, (size_t)rows * mXs * es));
  LSW_HIP(devMalloc(&dG, (size_t)rows * mGs * es));
  LSW_HIP(devMalloc(&dT, (size_t)nrhs * mTraceEntries * es));
  // (the final gather writes the output directly: no DD accumulators)
  if (!mGatherFinal) LSW_HIP(devMalloc(&dY, (size_t)4 * nrhs * N() * sizeof(double2)));
  // the solve kernels read (and multiply by zero) entries of X, R and G they
  // do not use: they must hold finite values from the start (cleared on the
  // problem's stream, ahead of its kernels)
  LSW_HIP(hipMemsetAsync(dX, 0, (size_t)rows * mXs * es, mStream));
  LSW_HIP(hipMemsetAsync(dXs, 0, (size_t)rows * mXs * es, mStream));
  LSW_HIP(hipMemsetAsync(dR, 0, (size_t)rows * mXs * es, mStream));
  LSW_HIP(hipMemsetAsync(dG, 0, (size_t)rows * mGs * es, mStream));
  // F is the final solves' right-hand-side buffer: stage 1 rewrites every
  // point of every used lane in each application
  LSW_HIP(hipMemsetAsync(dF, 0, (size_t)rows * mXs * es, mStream));
  // The solves may run on a caller's stream that is not ordered after the
  // DD's own (non-blocking) stream: such a call waits for the clears first
  // (ownWorkDone).  Buffers only grow, so this happens a few times per DD.
  markOwnWork();
  mBufRhs = nrhs;
}

void DDCheckerboard::markOwnWork() const {
  if (!mOwnEvent) LSW_HIP(hipEventCreateWithFlags(&mOwnEvent, hipEventDisableTiming));
  LSW_HIP(hipEventRecord(mOwnEvent, mStream));
  mOwnPending = true;
}

void DDCheckerboard::ownWorkDone(hipStream_t s) const {
  if (s != mStream && mOwnPending) {
    LSW_HIP(hipEventSynchronize(mOwnEvent));
    mOwnPending = false;
  }
}

// Rectangles of local grid points where the right-hand side of a slot can be
// nonzero: the windowed owned set (stage 1) or the two trace sets.
std::vector<Rect> DDCheckerboard::slotSupport(const SlotDev& s) const {
  std::vector<Rect> out;
  const CellGeom& c = mCells[s.cell];
  if (s.kind == 0) {
    const int slr = s.d & 1, sud = s.d >> 1;
    const bool wantColTrace = (s.d == 1 || s.d == 3), wantRowTrace = (s.d == 2 || s.d == 3);
    std::vector<std::pair<int, int>> cr, rr;
    for (int i = c.vlr[slr][2]; i < c.vlr[slr][5]; ++i) {
      if ((mColT[i + c.ox] != 0) != wantColTrace) continue;
      if (!cr.empty() && cr.back().second == i) cr.back().second = i + 1;
      else cr.push_back({i, i + 1});
    }
    for (int j = c.vud[sud][2]; j < c.vud[sud][5]; ++j) {
      if ((mRowT[j + c.oy] != 0) != wantRowTrace) continue;
      if (!rr.empty() && rr.back().second == j) rr.back().second = j + 1;
      else rr.push_back({j, j + 1});
    }
    for (auto& a : cr)
      for (auto& b : rr) out.push_back({a.first, a.second, b.first, b.second});
    if (out.empty()) out.push_back({0, 0, 0, 0});  // zero right-hand side: visit nothing
    return out;
  }
  for (int k = 0; k < 2; ++k) {
    if (s.setCol[k] >= 0 && s.setColHi[k] > s.setColLo[k])
      out.push_back({s.setCol[k], s.setCol[k] + 1, s.setColLo[k], s.setColHi[k]});
    if (s.setRow[k] >= 0 && s.setRowHi[k] > s.setRowLo[k])
      out.push_back({s.setRowLo[k], s.setRowHi[k], s.setRow[k], s.setRow[k] + 1});
  }
  if (out.empty()) out.push_back({0, 0, 0, 0});
  return out;
}

void DDCheckerboard::allocBuffers() {
  mXs = mFactors->maxDofs();
  mGs = std::max<long long>(mFactors->maxG(), 1);
  ensureBuffers(1);
}

void DDCheckerboard::deviceBytesEstimate(int precision, int nrhs, double* steady, double* peak,
                                        std::vector<std::shared_ptr<const NdTree>>* keep) const {
  LSW_REQUIRE(!mCells.empty(), LSW_ERR_STATE, "deviceBytesEstimate before setUp");
  std::map<std::pair<int, int>, std::shared_ptr<const NdTree>> trees;
  double entries = 0, schur = 0, local = 0;
  std::vector<double> wsPerHeight(kMaxHeight + 1, 0.0);
  for (const auto& c : mCells) {
    auto& t = trees[{c.nx, c.ny}];
    if (!t) {
      t = sharedTree(c.nx, c.ny);
      if (keep) keep->push_back(t);
    }
    entries += (double)t->factorSize;
    schur += (double)t->schurSize;
    local += (double)c.nx * c.ny;
    for (const auto& nd : t->nodes)
      if (nd.height <= kMaxHeight) wsPerHeight[nd.height] += (double)(nd.e + nd.b) * (2 * nd.e + nd.b);
  }
  const double es = precision == LSW_FP32 ? 8.0 : 16.0;
  const double n = (double)N();
  const double r = std::max(1, nrhs);
  const double nc = (double)mCells.size();
  double maxDofs = 0, maxG = 0, tmax = 0;
  for (const auto& c : mCells) {
    maxDofs = std::max(maxDofs, (double)c.nx * c.ny);
    maxG = std::max(maxG, (double)trees[{c.nx, c.ny}]->gSize);
    tmax = std::max(tmax, (double)std::max(c.nx, c.ny));
  }
  // stencils (local + global), squared slowness
  const double fixed = 80.0 * local + 8.0 * local + 80.0 * n;
  // sweep buffers per right-hand side: X, Xs and R (4 x cells slots), G, the
  // summed sources F, traces, and 4 FP64 accumulators of global vectors
  const double slots = 4.0 * nc;
  const double perRhs = es * (3.0 * slots * maxDofs + slots * maxG + 4.0 * nc * maxDofs + 4.0 * nc * 72.0 * tmax) +
                        64.0 * n;
  // solve plans: about 32 bytes per local DOF per block width, a dozen widths
  const double plans = 12.0 * 32.0 * local;
  // allocation granularity: the 4 MiB descriptor blocks of the DD's and the
  // factor set's arenas and the 2 MiB pages of every larger buffer (measured:
  // up to 32 MB over the byte counts above on cells of ~16k points)
  const double grain = 40.0 * (1 << 20);
  const double s = entries * es + fixed + r * perRhs + plans + grain;
  double ws = 0;
  for (double w : wsPerHeight) ws = std::max(ws, w);
  // fronts, Schur complements and factors while factorising: FP32 work
  // stores its factors directly; FP64 work for FP32 solves adds the copy
  const double wes = factorWorkFp32(precision) ? 8.0 : 16.0;
  ws = std::min(ws, (double)kFactorWsEntries) * wes;
  const double f = fixed + perRhs + (entries + schur) * wes + ws +
                   (precision == LSW_FP32 && wes == 16.0 ? entries * 8.0 : 0.0) + grain;
  if (steady) *steady = s;
  if (peak) *peak = std::max(s, f);
}

void DDCheckerboard::factorize() {
  LSW_REQUIRE(mFactors != nullptr, LSW_ERR_STATE, "setUp must be called before factorize");
  if (!mFactors->factorized()) mFactors->factorize(mPrecision, mStream);
}

// ------------------------------------------------------------- operations

void DDCheckerboard::applyPrecondDev(const double2* x, double2* y, hipStream_t s, int nrhs, bool replay) const {
  s = useStream(s);
  if (!mUseGraphs || !replay) {
    applyPrecondImpl(x, y, s, nullptr, nrhs);
    return;
  }
  // One M^{-1} application is ~30 launches per sweep step.  The sequence for
  // a given (x, y, nrhs) is captured once into a graph and replayed, which
  // takes the host launch cost off small problems.  Plans and buffers are
  // prepared eagerly first: nothing inside the capture allocates or waits.
  LSW_REQUIRE(mFactors && mFactors->factorized(), LSW_ERR_STATE, "applyPrecond before factorize");
  LSW_REQUIRE(nrhs >= 1, LSW_ERR_ARG, "applyPrecond: nrhs must be positive");
  const int before = mBufRhs;
  ensureBuffers(nrhs);
  ownWorkDone(s);  // (a host wait, so not inside the capture)
  if (mBufRhs != before) clearGraphs();  // captured graphs point at the old buffers
  for (const auto& st : mSteps) (void)planFor(st, nrhs);
  const GraphKey key{x, y, nrhs, s};
  auto it = mGraphs.find(key);
  if (it == mGraphs.end()) {
    if (mGraphs.size() >= kMaxGraphs) clearGraphs();
    hipGraph_t g = nullptr;
    LSW_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      applyPrecondImpl(x, y, s, nullptr, nrhs);
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    LSW_HIP(hipStreamEndCapture(s, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    LSW_HIP(e);
    it = mGraphs.emplace(key, ex).first;
  }
  LSW_HIP(hipGraphLaunch(it->second, s));
}

const PlanShape& DDCheckerboard::shapeFor(int nrhs) const {
  auto it = mShapes.find(nrhs);
  if (it == mShapes.end()) it = mShapes.emplace(nrhs, PlanShape::defaults()).first;
  return it->second;
}

void DDCheckerboard::dropPlans(int nrhs) const {
  clearGraphs();  // captured graphs point at the plans' work lists
  for (const auto& st : mSteps) st.plans.erase(nrhs);
}

// Launch-shape autotuning: time M^{-1} for blocks of nrhs right-hand sides
// (random input, 1 warm-up + 3 timed applications each) under the default
// PlanShape, then change one choice at a time (column-split threshold and
// row limit of latency-bound launches, column-split budget and cap of the
// launches of few large entries, matrix cores) and keep a change when it is at least 1%
// faster.  The plans for nrhs are rebuilt with the winner on next use.
DDCheckerboard::TuneResult DDCheckerboard::autotune(int nrhs) const {
  LSW_REQUIRE(mFactors && mFactors->factorized(), LSW_ERR_STATE, "autotune before factorize");
  LSW_REQUIRE(nrhs >= 1, LSW_ERR_ARG, "autotune: nrhs must be positive");
  ensureBuffers(nrhs);
  const long long n = N() * nrhs;
  std::vector<double2> hx((size_t)n);
  unsigned long long st = 0x9E3779B97F4A7C15ULL;
  for (auto& v : hx) {  // fixed pseudo-random input (xorshift)
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    const double a = (double)(st >> 11) / 9007199254740992.0;
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    const double b = (double)(st >> 11) / 9007199254740992.0;
    v = make_double2(a - 0.5, b - 0.5);
  }
  double2 *dx = nullptr, *dy = nullptr;
  LSW_HIP(devMalloc(&dx, (size_t)n * sizeof(double2)));
  LSW_HIP(devMalloc(&dy, (size_t)n * sizeof(double2)));
  LSW_HIP(uploadSync(dx, hx.data(), (size_t)n * sizeof(double2)));
  hipEvent_t e0, e1;
  LSW_HIP(hipEventCreate(&e0));
  LSW_HIP(hipEventCreate(&e1));
  auto timeShape = [&](const PlanShape& sh) {
    mShapes[nrhs] = sh;
    dropPlans(nrhs);
    applyPrecondImpl(dx, dy, mStream, nullptr, nrhs);  // builds the plans, warms the caches
    LSW_HIP(hipEventRecord(e0, mStream));
    for (int k = 0; k < 3; ++k) applyPrecondImpl(dx, dy, mStream, nullptr, nrhs);
    LSW_HIP(hipEventRecord(e1, mStream));
    LSW_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LSW_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms / 3.0;
  };
  TuneResult r;
  PlanShape best = PlanShape::defaults();
  r.defaultMs = r.bestMs = timeShape(best);
  r.tried = 1;
  auto consider = [&](PlanShape cand) {
    const double ms = timeShape(cand);
    ++r.tried;
    if (ms < 0.99 * r.bestMs) {
      r.bestMs = ms;
      best = cand;
    }
  };
  for (long long lw : {0LL, 1024LL, 16384LL}) {
    PlanShape c = best;
    c.latencyWaves = lw;
    consider(c);
  }
  for (int lr : {32, 64}) {
    PlanShape c = best;
    c.latencyRows = lr;
    consider(c);
  }
  for (long long wb : {0LL, 16384LL}) {
    PlanShape c = best;
    c.wideBudget = wb;
    consider(c);
  }
  for (int cap : {4, 16}) {  // 16: FP32 tasks of up to 4 right-hand sides (clamped to 8 otherwise)
    if (cap == 16 && !(mPrecision == LSW_FP32 && nrhs == 1)) continue;
    PlanShape c = best;
    c.wideCap = cap;
    consider(c);
  }
  if (mPrecision == LSW_FP32 && nrhs > 1) {
    PlanShape c = best;
    c.mfma = !c.mfma;
    consider(c);
    c = best;
    c.mfmaWideCols = c.mfmaWideCols >= (1 << 30) ? PlanShape().mfmaWideCols : (1 << 30);
    consider(c);
    c = best;
    c.mfmaPair = !c.mfmaPair;
    consider(c);
    c = best;
    c.mfmaGroup = !c.mfmaGroup;
    consider(c);
    if (best.mfmaGroup)
      for (int rows : {0, 16, 32}) {
        c = best;
        c.grpMinRowsFwd = c.grpMinRowsBwd = rows;
        consider(c);
      }
  }
  mShapes[nrhs] = best;
  dropPlans(nrhs);
  r.shape = best;
  LSW_HIP(hipEventDestroy(e0));
  LSW_HIP(hipEventDestroy(e1));
  (void)devFree(dx);
  (void)devFree(dy);
  return r;
}

void DDCheckerboard::clearGraphs() const {
  if (mGraphs.empty()) return;
  hipStream_t last = nullptr;
  for (auto& kv : mGraphs)
    if (kv.first.s != last) (void)hipStreamSynchronize(last = kv.first.s);
  for (auto& kv : mGraphs) (void)hipGraphExecDestroy(kv.second);
  mGraphs.clear();
}

void DDCheckerboard::applyPrecondProfiled(const double2* x, double2* y, SolveProfile& prof, int nrhs) const {
  applyPrecondImpl(x, y, mStream, &prof, nrhs);
  LSW_HIP(hipStreamSynchronize(mStream));
}


void DDCheckerboard::applyPrecondImpl(const double2* x, double2* y, hipStream_t s, SolveProfile* prof,
                                      int nrhs) const {
  LSW_REQUIRE(mFactors && mFactors->factorized(), LSW_ERR_STATE, "applyPrecond before factorize");
  LSW_REQUIRE(nrhs >= 1, LSW_ERR_ARG, "applyPrecond: nrhs must be positive");
  LSW_REQUIRE((long long)mMaxSlots * nrhs <= 65535, LSW_ERR_CAPACITY, "too many right-hand sides in one block");
  s = useStream(s);
  ensureBuffers(nrhs);
  ownWorkDone(s);
  const size_t es = (mPrecision == LSW_FP64) ? sizeof(double2) : sizeof(float2);
  const long long n = N();
  auto zero = [&](void* p, long long words) {
    if (words > 0)
      hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<long long>((words + 255) / 256, 8192)), dim3(256), 0,
                         s, (unsigned long long*)p, words);
  };
  if (!mOwnedPartition) zero(dY, 4LL * nrhs * n * 2);
  zero(dT, (long long)(nrhs * mTraceEntries * es / 8));
  const CellDev* dCells = mFactors->deviceCells();
  const double2* dCoef = mFactors->deviceCoef();
  // F, the final solves' right-hand sides, is laid out by the final step's
  // plan; a (d, cell) group's final slot is the group index
  const Step& stF = mSteps.back();
  const SolvePlan& planF = planFor(stF, nrhs);
  LSW_REQUIRE(stF.deferred && planF.taskRows() <= mTaskRows, LSW_ERR_STATE, "final step plan does not fit F");
  LSW_REQUIRE(planF.slotTQ.size() >= (size_t)nrhs * stF.nSlots, LSW_ERR_STATE, "final plan misses slots");
  for (size_t k = 0; k < (size_t)nrhs * stF.nSlots; ++k)
    LSW_REQUIRE(planF.slotTQ[k] >= 0, LSW_ERR_STATE, "a final solve has no task");
  const int tb = 256;
  const unsigned gxT = (unsigned)((mTmax + 63) / 64);
  // per sweep step: right-hand sides, their sum per (d, cell), the solves
  // (backward only where traces are read), trace extraction and transfer;
  // the final step solves each (d, cell) once in full from F and combines
  bool s1InF = false;
  auto run = [&](auto tag) {
    using T2 = decltype(tag);
    for (size_t si = 0; si < mSteps.size(); ++si) {
      const Step& st = mSteps[si];
      const unsigned ns = (unsigned)st.nSlots * nrhs, ng = (unsigned)st.nGroups * nrhs;
      const unsigned nc = (unsigned)st.nCopies * nrhs;
      // stage 1 writes whole buffers (X and F); the sweep steps write only
      // their trace sets into X, which is zero elsewhere (cleared once here
      // after stage 1, then by each sweep step's k_post), and add them into F
      const bool sparse = si > 0 && !st.deferred;
      // the step's local vectors are laid out by its plan's tasks
      const SolvePlan& plan = planFor(st, nrhs);
      LSW_REQUIRE(plan.taskRows() <= mTaskRows, LSW_ERR_STATE, "solve scratch smaller than the plan needs");
      LSW_REQUIRE((long long)plan.slotTQ.size() >= (long long)ns, LSW_ERR_STATE, "plan does not cover the step's slots");
      const int W = plan.width;
      const unsigned nt = (unsigned)plan.ntasks;
      const unsigned gxP = (unsigned)std::min<long long>((mFactors->maxDofs() + tb - 1) / tb, 1024);
      // kernels instantiated per task width (1, 4, 8)
      auto withWidth = [&](int w, auto&& f) {
        if (w == 1)
          f(std::integral_constant<int, 1>{});
        else if (w == kTaskRhs4)
          f(std::integral_constant<int, kTaskRhs4>{});
        else
          f(std::integral_constant<int, kMaxRhs>{});
      };
      // stage 1 in F's layout writes and solves from F alone; X then stays
      // zero off the sweep steps' trace sets and needs no clearing
      const bool inF = si == 0 && mStage1InF && plan.width == planF.width && plan.slotTQ == planF.slotTQ;
      if (si == 0) s1InF = inF;
      if (si == 1 && !s1InF) zero(dX, (long long)((size_t)mTaskRows * mXs * es / 8));
      if (sparse) {
        // (one right-hand side: few, latency-bound groups take 1024 threads;
        // blocks have nrhs times more groups, which fill the chip at 256)
        hipLaunchKernelGGL(k_build_rhs_groups<T2>, dim3(ng), dim3(nrhs == 1 ? 1024 : 256), 0, s, st.dGroups, st.nGroups, dCells,
                           st.dSlots, st.nSlots, plan.dSlotTQ, W, mTraceEntries, dCoef, (const T2*)dT, (T2*)dX, mXs,
                           (T2*)dF, planF.dSlotTQ, planF.width, stF.nSlots);
      } else if (!st.deferred) {
        auto build = [&](auto w) {
          hipLaunchKernelGGL((k_build_rhs_tasks<T2, decltype(w)::value>), dim3(gxP, nt), dim3(tb), 0, s, dCells,
                             st.dSlots, st.nSlots, plan.dTasks, mTraceEntries, n, x, dColTrace, dRowTrace, mGy, dCoef,
                             (const T2*)dT, (T2*)dX, mXs, (T2*)dF, planF.dSlotTQ, planF.width, stF.nSlots,
                             inF ? 0 : 1);
        };
        withWidth(W, build);
      }
      // solutions go to dXs: dX keeps the right-hand sides; the final solves
      // read F (stage 1's right-hand sides plus the sweep sums, nothing to build)
      mFactors->solve(plan, st.deferred || inF ? dF : dX, dXs, dR, dG, mXs, mGs, s, prof);
      if (st.deferred && mGatherFinal) {
        const size_t need = (size_t)nrhs * st.nSlots;
        LSW_REQUIRE(plan.slotTQ.size() >= need, LSW_ERR_STATE, "final gather: slot map shorter than the step");
        for (size_t k = 0; k < need; ++k)
          LSW_REQUIRE(plan.slotTQ[k] >= 0, LSW_ERR_STATE, "final gather: a final solve has no task");
        auto gat = [&](auto w) {
          hipLaunchKernelGGL((k_gather_final<T2, decltype(w)::value>),
                             dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)), dim3(256), 0, s, dCells,
                             dOwn, plan.dSlotTQ, mXs, st.nSlots, (int)mCells.size(), mNC, (const T2*)dXs, y, mGx,
                             mGy, nrhs);
        };
        withWidth(W, gat);
      } else if (st.deferred) {
        auto acc = [&](auto w) {
          hipLaunchKernelGGL((k_accumulate_tasks<T2, decltype(w)::value>), dim3(gxP, nt), dim3(tb), 0, s, dCells,
                             st.dSlots, st.nSlots, plan.dTasks, (const T2*)dXs, mXs, dY, n, mGy,
                             mOwnedPartition ? 1 : 0);
        };
        withWidth(W, acc);
      } else {
        hipLaunchKernelGGL(k_post<T2>, dim3(gxT, ns), dim3(64), 0, s, dCells, st.dSlots, st.nSlots, plan.dSlotTQ, W,
                           mTraceEntries, (const T2*)dXs, (T2*)dX, mXs, (T2*)dT);
        if (st.nCopies)
          hipLaunchKernelGGL(k_copy_traces<T2>, dim3(gxT, nc), dim3(64), 0, s, st.dCopies, st.nCopies,
                             mTraceEntries, (T2*)dT, mTmax);
      }
    }
  };
  if (mPrecision == LSW_FP64)
    run(double2{});
  else
    run(float2{});
  if (!mGatherFinal)
    hipLaunchKernelGGL(k_combine, dim3((unsigned)std::min<long long>((n * nrhs + 255) / 256, 8192)), dim3(256), 0,
                       s, dY, y, n, nrhs);
  LSW_HIP(hipGetLastError());
}

void DDCheckerboard::applyADev(const double2* x, double2* y, hipStream_t s) const {
  applyADevBlock(x, y, 1, s);
}

void DDCheckerboard::applyADevBlock(const double2* x, double2* y, int nrhs, hipStream_t s) const {
  LSW_REQUIRE(dGlobalCoef != nullptr, LSW_ERR_STATE, "DD has no device state");
  s = useStream(s);
  ownWorkDone(s);
  const long long n = N();
  if (nrhs > 0)
    hipLaunchKernelGGL(k_matvec, dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)), dim3(256), 0, s,
                       dGlobalCoef, x, y, mGx, mGy, nrhs);
  LSW_HIP(hipGetLastError());
}

void DDCheckerboard::assembleRHSDev(const std::vector<DoubleVector>& srcs, double2* b, hipStream_t s) const {
  LSW_REQUIRE(deviceReady(), LSW_ERR_STATE, "DD has no device state");
  s = useStream(s);
  std::vector<double> xy;
  for (auto& p : srcs) {
    xy.push_back(p[0]);
    xy.push_back(p[1]);
  }
  double* dsrc = nullptr;
  LSW_HIP(devMalloc(&dsrc, std::max<size_t>(1, xy.size()) * sizeof(double)));
  LSW_HIP(hipMemcpyAsync(dsrc, xy.data(), xy.size() * sizeof(double), hipMemcpyHostToDevice, s));
  const long long n = N();
  hipLaunchKernelGGL(k_point_sources, dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)), dim3(256), 0,
                     s, dsrc, (int)srcs.size(), mMeshSize, mGlobal.asmb.x0(), mGlobal.asmb.hx(), mGlobal.asmb.y0(),
                     mGlobal.asmb.hy(), mGx, mGy, b);
  LSW_HIP(hipGetLastError());
  LSW_HIP(hipStreamSynchronize(s));
  IdleFrees idle;  // only the kernel above, done, read it
  (void)devFree(dsrc);
}

void DDCheckerboard::ownedGlobalIndices(int c, std::vector<long long>& idx) const {
  const CellGeom& cg = mCells[c];
  idx.clear();
  for (int i = cg.vlr[0][2]; i < cg.vlr[0][5]; ++i)
    for (int j = cg.vud[0][2]; j < cg.vud[0][5]; ++j)
      idx.push_back((long long)(i + cg.ox) * mGy + (j + cg.oy));
}

void DDCheckerboard::gatherOwned(const std::vector<Complex>& g, std::vector<ComplexVector>& res) const {
  res.resize(mCells.size());
  std::vector<long long> idx;
  for (size_t c = 0; c < mCells.size(); ++c) {
    ownedGlobalIndices((int)c, idx);
    res[c].resize(idx.size());
    for (size_t k = 0; k < idx.size(); ++k) res[c][k] = g[idx[k]];
  }
}

void DDCheckerboard::scatterOwned(const std::vector<ComplexVector>& x, std::vector<Complex>& g) const {
  LSW_REQUIRE(x.size() == mCells.size(), LSW_ERR_ARG, "wrong number of cell vectors");
  g.assign((size_t)N(), 0.0);
  std::vector<long long> idx;
  for (size_t c = 0; c < mCells.size(); ++c) {
    ownedGlobalIndices((int)c, idx);
    LSW_REQUIRE(x[c].size() == idx.size(), LSW_ERR_ARG, "wrong owned vector length");
    for (size_t k = 0; k < idx.size(); ++k) g[idx[k]] = x[c][k];
  }
}

namespace {
struct DevBuf {
  double2* p = nullptr;
  explicit DevBuf(long long n) { LSW_HIP(devMalloc(&p, std::max<long long>(n, 1) * sizeof(double2))); }
  ~DevBuf() {
    if (p) (void)devFree(p);
  }
};
}  // namespace

void DDCheckerboard::applyPrecond(const std::vector<ComplexVector>& x, std::vector<ComplexVector>& res) const {
  std::vector<Complex> g;
  scatterOwned(x, g);
  DevBuf dx(N()), dy(N());
  LSW_HIP(uploadSync(dx.p, g.data(), N() * sizeof(double2)));
  applyPrecondDev(dx.p, dy.p, mStream, 1, false);
  LSW_HIP(hipStreamSynchronize(mStream));
  LSW_HIP(hipMemcpy(g.data(), dy.p, N() * sizeof(double2), hipMemcpyDeviceToHost));
  gatherOwned(g, res);
}

void DDCheckerboard::applySystemMatrix(const std::vector<ComplexVector>& x, std::vector<ComplexVector>& res) const {
  std::vector<Complex> g;
  scatterOwned(x, g);
  DevBuf dx(N()), dy(N());
  LSW_HIP(uploadSync(dx.p, g.data(), N() * sizeof(double2)));
  applyADev(dx.p, dy.p, mStream);
  LSW_HIP(hipStreamSynchronize(mStream));
  LSW_HIP(hipMemcpy(g.data(), dy.p, N() * sizeof(double2), hipMemcpyDeviceToHost));
  gatherOwned(g, res);
}

void DDCheckerboard::assembleRHS(const RHS* rhs, std::vector<ComplexVector>& res) const {
  const auto* sp = dynamic_cast<const RHSSrcPts*>(rhs);
  LSW_REQUIRE(sp != nullptr, LSW_ERR_ARG, "assembleRHS supports RHSSrcPts");
  DevBuf db(N());
  assembleRHSDev(sp->getSrcPts(), db.p, mStream);
  std::vector<Complex> g((size_t)N());
  LSW_HIP(hipMemcpy(g.data(), db.p, N() * sizeof(double2), hipMemcpyDeviceToHost));
  gatherOwned(g, res);
}

}  // namespace lsw
