// Batched nested-dissection solves: the hot loop of the L-sweeps.
//
// One solve of a cell is, per tree height, a set of dense matrix-vector
// products (forward: g_p = W_p r_E + children updates; backward:
// x_E = [F11^{-1} | -F11^{-1} F12] [r_E; x_B]).  A launch covers one height of
// every task of a sweep step.  Work lists are flat {task, node, row chunk}
// entries prepared on the host once per step:
//   * the forward list only holds nodes whose subtree region meets the
//     right-hand side support (sweep right-hand sides live on 2-4 grid lines);
//     a root has no boundary rows, so its forward entry only stores r_E;
//   * the backward list holds the nodes whose solution values are read (all
//     of them for a full solve), and skips the F11^{-1} r_E columns of nodes
//     the forward pass did not visit (their r_E is zero); a root of a single
//     row chunk forms its r_E from X and the children's updates while it
//     gathers, and has no forward entry.
// Everything an entry needs is one SolveEntry, one SolveTaskDev and one
// SolveNode (loaded side by side).  Sweep steps are latency-bound (a few
// microseconds of work per launch), so an entry's chain of dependent memory
// accesses is kept short: entry -> task/node -> index pools -> vector values
// -> products; the factor values, the children's updates a forward row adds
// and the destination index a backward row stores to are all loaded in the
// same rounds as the gather, not after the product.
// Two launch shapes:
//   * packed: a wavefront holds P = 64/RP entries side by side (RP lanes =
//     rows of one entry), `wpb` wavefronts per workgroup run independently,
//     each lane group walks `rounds` entries; used wherever a height has
//     enough entries (the many small fronts near the leaves);
//   * wide: W = 1..8 wavefronts split the columns of one entry (64/RP column
//     groups per wavefront) and reduce in LDS; used for heights with few,
//     large entries, and for every launch with few wavefronts (the launches
//     of a sweep step), where a packed lane walking its whole row would make
//     the product a long chain of memory round trips.
// Factors are column-major, so the lanes of an entry read consecutive rows of
// one column (coalesced); right-hand-side values are broadcast from LDS.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "lsw_cplx.h"
#include "lsw_factor.h"

namespace lsw {

namespace {
constexpr double kMaxLds = 65536.0;  // bytes of dynamic LDS per workgroup

// LDS accesses of one wavefront are performed in order; this only stops the
// compiler from moving accesses across the point.
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Every global load below is issued unconditionally, from an address that
// is valid for every lane (indices are clamped, unused right-hand sides read
// slot 0, a missing child reads offset 0 of the update buffer), and every
// loaded value feeds arithmetic on every path: values a lane must ignore are
// multiplied by zero (cmask), LDS staging is written at clamped positions
// (rewriting a value already there), and a store a lane must not make goes
// to a trash word.  A load whose value is only used under a divergent
// branch is moved into that branch by the compiler and waited for before
// the branches join, which turns a chain of independent loads into one
// memory round trip each; a sweep step's launches move only a few MB, so
// their duration is that chain.  Branches that skip loads test wave-uniform
// conditions (waveAny).  The update buffers are zeroed when allocated, so a
// masked read never meets a NaN.
constexpr int kDotDepth = 8;  // factor loads per batch of a product

// A packed launch of more than this many wavefronts gives each lane group
// several entries in turn ("rounds").  By default every lane group takes one
// entry: on the 1600^2 lens, rounds that kept 4096 wavefronts were 4% slower
// overall than one entry per lane group (LSW_PACKED_WAVES overrides it).
long long packedWaveTarget() {
  static const long long v = [] {
    const char* e = std::getenv("LSW_PACKED_WAVES");
    return e ? std::max(1LL, std::atoll(e)) : (1LL << 40);
  }();
  return v;
}


// Workgroups are dispatched to the 8 XCDs (each with its own L2) in turn:
// workgroup b runs on XCD b % 8.  This gives each XCD a contiguous range of
// the work, so that neighbouring entries (the tasks of one cell reading the
// same factor block) meet in one L2.  A bijection of [0, G).
__device__ __forceinline__ unsigned xcdRemap(unsigned b, unsigned G) {
  constexpr unsigned kXcd = 8;
  const unsigned q = G / kXcd, r = G % kXcd, x = b % kXcd;
  return x * q + (x < r ? x : r) + b / kXcd;
}

// true in every lane when pred holds in any lane of the wavefront
__device__ __forceinline__ bool waveAny(bool pred) { return __ballot(pred) != 0; }

// values a lane gathers per batch: all index loads of a batch, then all
// value loads
template <int NR>
constexpr int gatherBatch() {
  return NR >= 8 ? 2 : NR >= 4 ? 4 : 8;
}

// x if keep else 0, by arithmetic (x finite)
template <typename T2>
__device__ __forceinline__ T2 cmask(T2 x, bool keep) {
  const auto s = keep ? decltype(x.x)(1) : decltype(x.x)(0);
  T2 y;
  y.x = x.x * s;
  y.y = x.y * s;
  return y;
}

// store target of a lane that must not store
__device__ double2 gSolveTrash[512];

template <typename T2>
__device__ __forceinline__ T2* trashSlot() {
  return reinterpret_cast<T2*>(gSolveTrash) + (threadIdx.x & 511);
}

// store target of NR <= 8 consecutive values a lane must not store
template <typename T2, int NR>
__device__ __forceinline__ T2* trashRow() {
  static_assert(NR <= 8, "trash rows hold at most 8 values");
  return reinterpret_cast<T2*>(gSolveTrash) + (threadIdx.x & 63) * NR;
}

// m[u] = M[(k + u*kstep)*ld] for k + u*kstep < K, else M[0] (which must be
// valid; it meets the zero pad column in rowAcc)
template <typename T2, int D>
__device__ __forceinline__ void rowLoad(const T2* __restrict__ M, long long ld, int k, int K, int kstep, T2 (&m)[D]) {
#pragma unroll
  for (int u = 0; u < D; ++u) {
    const int kk = k + u * kstep;
    m[u] = M[(long long)(kk < K ? kk : 0) * ld];
  }
}

// acc[q] += sum_u m[u] * v[(k + u*kstep)*NR + q] over the terms below K:
// past K the product reads the staging's zero pad column instead (m is then
// finite but meaningless), so no select is needed per term.  The staging is
// column-major over right-hand sides: the NR values of one column are
// adjacent (wide LDS reads).  With many right-hand sides the values of one
// column are consumed before the next column's are read, which bounds the
// registers they hold.
template <typename T2, int NR, int D>
__device__ __forceinline__ void rowAcc(const T2 (&m)[D], int k, int K, int kstep, const T2* v, int pad,
                                       T2 (&acc)[NR]) {
#pragma unroll
  for (int u = 0; u < D; ++u) {
    const int kk = k + u * kstep;
    const T2* vk = v + (kk < K ? kk : pad) * NR;
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = cfma(m[u], vk[q], acc[q]);
    if constexpr (NR >= 8) __builtin_amdgcn_sched_barrier(0);
  }
}

// acc[q] += sum_{k = k0 + j*kstep < K} M[k*ld] * v[k*NR + q]: the first batch
// comes from pre (loaded before the gather, so that its latency overlaps
// it), the rest in batches of D, each loaded while the previous one is
// multiplied (one memory round trip per row instead of one per batch).  The
// loop runs while any lane of the wavefront has columns left, so its loads
// are not under a divergent branch; lanes past their K load M[0] and meet
// the zero pad.
template <typename T2, int NR, int D>
__device__ __forceinline__ void rowDotPre(const T2* __restrict__ M, long long ld, int k0, int K, int kstep,
                                          const T2* v, int pad, const T2 (&pre)[D], T2 (&acc)[NR]) {
  T2 cur[D];
#pragma unroll
  for (int u = 0; u < D; ++u) cur[u] = pre[u];
  for (int k = k0;; k += D * kstep) {
    const int kn = k + D * kstep;
    const bool more = waveAny(kn < K);
    T2 nxt[D];
    if (more) rowLoad<T2, D>(M, ld, kn, K, kstep, nxt);
    rowAcc<T2, NR, D>(cur, k, K, kstep, v, pad, acc);
    if (!more) break;
#pragma unroll
    for (int u = 0; u < D; ++u) cur[u] = nxt[u];
  }
}

struct Item {
  SolveTaskDev tk;
  SolveNode nd;
  int task, r0, rows, flags;
};

// The update vectors G and the kept r_E (R) of a plan are task-major with the
// NR right-hand sides of a task interleaved: value q of position i of task t
// is at ((t * stride + i) * NR + q), so one position of all right-hand sides
// is one contiguous NR-value load or store.
template <typename T2, int NR>
__device__ __forceinline__ long long tix(const Item& it, long long stride, long long i) {
  return ((long long)it.task * stride + i) * NR;
}

template <typename T2, int NR>
__device__ __forceinline__ void loadNR(const T2* __restrict__ p, T2 (&o)[NR]) {
  const T2* a = reinterpret_cast<const T2*>(__builtin_assume_aligned(p, NR * sizeof(T2) >= 16 ? 16 : sizeof(T2)));
#pragma unroll
  for (int q = 0; q < NR; ++q) o[q] = a[q];
}

template <typename T2, int NR>
__device__ __forceinline__ void storeNR(T2* __restrict__ p, const T2 (&o)[NR]) {
  T2* a = reinterpret_cast<T2*>(__builtin_assume_aligned(p, NR * sizeof(T2) >= 16 ? 16 : sizeof(T2)));
#pragma unroll
  for (int q = 0; q < NR; ++q) a[q] = o[q];
}

__device__ __forceinline__ Item loadItem(const SolveEntry* __restrict__ ent, const SolveTaskDev* __restrict__ tasks,
                                         const SolveNode* __restrict__ nodes, long long idx) {
  Item it;
  const SolveEntry en = ent[idx];
  it.tk = tasks[en.task];
  it.nd = nodes[en.node];
  it.task = en.task;
  it.r0 = en.r0rows & 0xffff;
  it.rows = en.r0rows >> 16;
  it.flags = en.flags;
  return it;
}

// r_E of a front: v[k*NR + q] = X[E[k]] - g_c1[m1[k]] - g_c2[m2[k]] for the
// lane's k in [k0, K) (step kstep) when `act`; with keepR also kept in Rb
// for the backward pass.
template <typename T2, int NR>
__device__ __forceinline__ void gatherFrontRhs(const Item& it, bool act, int k0, int K, int kstep,
                                               const int* __restrict__ E, const int* __restrict__ m1,
                                               const T2* __restrict__ X, T2* __restrict__ Rb, bool keepR,
                                               const T2* __restrict__ G, long long xs, long long gs, T2* v) {
  constexpr int B = gatherBatch<NR>();
  const int F = it.nd.e + it.nd.b;
  const bool c1 = (it.flags & kEntChild1) != 0, c2 = (it.flags & kEntChild2) != 0;
  const bool kids = waveAny(act && (c1 || c2));
  if (!act) K = k0;
  for (int kb = k0; kb < K; kb += B * kstep) {
    int dof[B], i1[B], i2[B];
    bool h1[B], h2[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int kk = kb + u * kstep;
      const int kc = kk < K ? kk : kb;
      dof[u] = E[kc];
      int a1 = -1, a2 = -1;
      if (kids) {
        a1 = m1[kc];
        a2 = m1[F + kc];
      }
      h1[u] = c1 && a1 >= 0;
      h2[u] = c2 && a2 >= 0;
      i1[u] = h1[u] ? it.nd.g1 + a1 : 0;
      i2[u] = h2[u] ? it.nd.g2 + a2 : 0;
    }
    // all loads of the batch first: the stores below may alias them as far
    // as the compiler knows
    T2 val[B][NR];
#pragma unroll
    for (int u = 0; u < B; ++u) loadNR<T2, NR>(X + tix<T2, NR>(it, xs, dof[u]), val[u]);
    if (kids) {
#pragma unroll
      for (int u = 0; u < B; ++u) {
        T2 g1v[NR], g2v[NR];
        loadNR<T2, NR>(G + tix<T2, NR>(it, gs, i1[u]), g1v);
        loadNR<T2, NR>(G + tix<T2, NR>(it, gs, i2[u]), g2v);
#pragma unroll
        for (int q = 0; q < NR; ++q) val[u][q] = csub(csub(val[u][q], cmask(g1v[q], h1[u])), cmask(g2v[q], h2[u]));
      }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int kk = kb + u * kstep;
      const int kc = kk < K ? kk : kb;
#pragma unroll
      for (int q = 0; q < NR; ++q) v[kc * NR + q] = val[u][q];  // past K: the value of kb again
      if (keepR && kk < K) storeNR<T2, NR>(Rb + tix<T2, NR>(it, xs, dof[u]), val[u]);
    }
  }
}

// forward gather: v[k*NR + q] = r_E[k]; the first row chunk also keeps r_E
// for the backward pass
template <typename T2, int NR>
__device__ __forceinline__ void fwdGather(const Item& it, bool act, int k0, int kstep, const int* __restrict__ poolE,
                                          const int* __restrict__ poolMap, const T2* __restrict__ X,
                                          T2* __restrict__ Rb, const T2* __restrict__ G, long long xs, long long gs,
                                          T2* v) {
  gatherFrontRhs<T2, NR>(it, act, k0, it.nd.e, kstep, poolE + it.nd.E, poolMap + it.nd.map, X, Rb, it.r0 == 0, G,
                         xs, gs, v);
}

// The children's updates a forward row adds to its product do not depend
// on the product: they are loaded with the gather and start the
// accumulators.
template <typename T2, int NR>
__device__ __forceinline__ void fwdRowAdd(const Item& it, bool ok, int r, const int* __restrict__ poolMap,
                                          const T2* __restrict__ G, long long gs, T2 (&acc)[NR]) {
  const int e = it.nd.e, F = e + it.nd.b;
  const int f = ok ? e + it.r0 + r : 0;
  const int* m1 = poolMap + it.nd.map;
  const bool c1 = (it.flags & kEntChild1) != 0, c2 = (it.flags & kEntChild2) != 0;
#pragma unroll
  for (int q = 0; q < NR; ++q) acc[q] = czero<T2>();
  if (!waveAny(ok && (c1 || c2))) return;
  const int a1 = m1[f], a2 = m1[F + f];
  const bool h1 = ok && c1 && a1 >= 0, h2 = ok && c2 && a2 >= 0;
  const int i1 = h1 ? it.nd.g1 + a1 : 0, i2 = h2 ? it.nd.g2 + a2 : 0;
  T2 g1v[NR], g2v[NR];
  loadNR<T2, NR>(G + tix<T2, NR>(it, gs, i1), g1v);
  loadNR<T2, NR>(G + tix<T2, NR>(it, gs, i2), g2v);
#pragma unroll
  for (int q = 0; q < NR; ++q) acc[q] = cadd(cmask(g1v[q], h1), cmask(g2v[q], h2));
}

// forward scatter of row r: g_p[r0+r] = product + children's updates (all
// NR right-hand sides of the task: unused ones hold finite values nobody
// reads into a result)
template <typename T2, int NR>
__device__ __forceinline__ void fwdStore(const Item& it, bool ok, int r, const T2 (&acc)[NR], T2* __restrict__ G,
                                         long long gs) {
  T2* dst = ok ? G + tix<T2, NR>(it, gs, it.nd.gOff + it.r0 + r) : trashRow<T2, NR>();
  storeNR<T2, NR>(dst, acc);
}

// backward gather: v[k*NR + q] = r_E[k] (k < e) or x_B[k-e] (from the
// solution buffer Xo).  The r_E of a root is formed here from the right-hand
// sides X and its children's updates when the plan says so (the solution
// goes to another buffer, or the root has one row chunk, whose single entry
// reads all of X[E] before it stores any of it); a root has no boundary, so
// its columns are all r_E.
template <typename T2, int NR>
__device__ __forceinline__ void bwdGather(const Item& it, bool act, int k0, int kstep, const int* __restrict__ poolE,
                                          const int* __restrict__ poolB, const int* __restrict__ poolMap,
                                          const T2* X, const T2* Xo, const T2* __restrict__ Rb,
                                          const T2* __restrict__ G, long long xs, long long gs, T2* v) {
  constexpr int B = gatherBatch<NR>();
  const int e = it.nd.e, F = e + it.nd.b;
  const bool root = (it.flags & kEntRootRhs) != 0;
  if (waveAny(act && root))
    gatherFrontRhs<T2, NR>(it, act && root, k0, e, kstep, poolE + it.nd.E, poolMap + it.nd.map, X, nullptr, false,
                           G, xs, gs, v);
  if (!waveAny(act && !root)) return;
  const int* E = poolE + it.nd.E;
  const int* Bp = poolB + it.nd.B;
  const int K = (act && !root) ? F : k0;
  for (int kb = k0; kb < K; kb += B * kstep) {
    int dof[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int kk = kb + u * kstep;
      const int kc = kk < K ? kk : kb;
      dof[u] = *(kc < e ? E + kc : Bp + (kc - e));
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int kk = kb + u * kstep;
      const int kc = kk < K ? kk : kb;
      // r_E from the task-major R, x_B from the right-hand sides' solutions
      T2 rv[NR], xb[NR];
      const bool fromR = kc < e;
      loadNR<T2, NR>(Rb + tix<T2, NR>(it, xs, fromR ? dof[u] : 0), rv);
      loadNR<T2, NR>(Xo + tix<T2, NR>(it, xs, dof[u]), xb);
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        T2 w;  // selected per component (a select of whole values went through scratch)
        w.x = fromR ? rv[q].x : xb[q].x;
        w.y = fromR ? rv[q].y : xb[q].y;
        v[kc * NR + q] = w;  // past K: the value of kb again
      }
    }
  }
}

// backward store of x_E[row] for all NR right-hand sides of the task (the
// task's lanes past its right-hand sides are its own: nobody reads them)
template <typename T2, int NR>
__device__ __forceinline__ void bwdStore(const Item& it, bool ok, int dof, const T2 (&acc)[NR], T2* __restrict__ X,
                                         long long xs) {
  storeNR<T2, NR>(ok ? X + tix<T2, NR>(it, xs, dof) : trashRow<T2, NR>(), acc);
}

// ---------------------------------------------------------------- packed
// LDS: per wavefront P regions of NR x cols values (the last column is the
// zero pad, cleared here and never overwritten).  Lanes past the last
// entry work on the last entry and store nothing.
template <typename T2, int NR, bool FWD>
__global__ __launch_bounds__(256) void k_nd_packed(const SolveTaskDev* __restrict__ tasks,
                                                   const SolveNode* __restrict__ nodes,
                                                   const SolveEntry* __restrict__ ent, int nent, int rounds,
                                                   int rpLog, int cols, const int* __restrict__ poolE,
                                                   const int* __restrict__ poolB, const int* __restrict__ poolMap,
                                                   const T2* __restrict__ factor, const T2* X, T2* Xo,
                                                   T2* __restrict__ Rb, T2* __restrict__ G, long long xs,
                                                   long long gs) {
  extern __shared__ char smem_raw[];
  constexpr int D = kDotDepth;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int RP = 1 << rpLog, P = 64 >> rpLog;
  const int grp = lane >> rpLog, r = lane & (RP - 1);
  T2* v = reinterpret_cast<T2*>(smem_raw) + ((size_t)wave * P + grp) * NR * cols;
  const int pad = cols - 1;
  if (r < NR) v[pad * NR + r] = czero<T2>();  // RP >= 8 >= NR
  const long long waveId = (long long)xcdRemap(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wave;
  for (int round = 0; round < rounds; ++round) {
    const long long idx = (waveId * rounds + round) * P + grp;
    const bool valid = idx < nent;
    const Item it = loadItem(ent, tasks, nodes, valid ? idx : nent - 1);
    const bool ok = valid && r < it.rows;
    const int e = it.nd.e, F = e + it.nd.b;
    const long long ld = FWD ? it.nd.b : e;
    const int k0 = FWD ? 0 : ((it.flags & kEntVisited) ? 0 : e);
    const int K = ok ? (FWD ? e : F) : k0;
    // a lane without columns reads the factor base (valid) instead of its row
    const T2* M = K > k0 ? factor + it.tk.facBase + (FWD ? it.nd.fwdOff : it.nd.bwdOff) + it.r0 + r : factor;
    T2 pre[D];
    rowLoad<T2, D>(M, ld, k0, K, 1, pre);
    T2 acc[NR];
    int dof = 0;
    if constexpr (FWD) {
      fwdRowAdd<T2, NR>(it, ok, r, poolMap, G, gs, acc);
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = czero<T2>();
      dof = poolE[it.nd.E + (ok ? it.r0 + r : 0)];
    }
    if constexpr (FWD)
      fwdGather<T2, NR>(it, valid, r, RP, poolE, poolMap, X, Rb, G, xs, gs, v);
    else
      bwdGather<T2, NR>(it, valid, k0 + r, RP, poolE, poolB, poolMap, X, Xo, Rb, G, xs, gs, v);
    waveSync();
    rowDotPre<T2, NR, D>(M, ld, k0, K, 1, v, pad, pre, acc);
    if constexpr (FWD)
      fwdStore<T2, NR>(it, ok, r, acc, G, gs);
    else
      bwdStore<T2, NR>(it, ok, dof, acc, Xo, xs);
    waveSync();
  }
}

// ------------------------------------------------------------------ wide
// W wavefronts per entry: 64 rows x W*(64/RP) column groups; partial sums
// reduced by shuffles inside a wavefront and in a fixed order across
// wavefronts through LDS.  LDS: NR x cols values (the last column is the
// zero pad), then W x 64 x NR partials.
template <typename T2, int NR, int W, bool FWD>
__global__ __launch_bounds__(64 * W) void k_nd_wide(const SolveTaskDev* __restrict__ tasks,
                                                    const SolveNode* __restrict__ nodes,
                                                    const SolveEntry* __restrict__ ent, int nent, int cols,
                                                    const int* __restrict__ poolE, const int* __restrict__ poolB,
                                                    const int* __restrict__ poolMap, const T2* __restrict__ factor,
                                                    const T2* X, T2* Xo, T2* __restrict__ Rb, T2* __restrict__ G,
                                                    long long xs, long long gs) {
  extern __shared__ char smem_raw[];
  constexpr int D = kDotDepth;
  T2* v = reinterpret_cast<T2*>(smem_raw);
  T2* red = v + (size_t)NR * cols;
  const int pad = cols - 1;
  if (threadIdx.x < NR) v[pad * NR + threadIdx.x] = czero<T2>();  // ordered by the barrier after the gather
  const long long idx = xcdRemap(blockIdx.x, gridDim.x);
  if (idx >= nent) return;
  const Item it = loadItem(ent, tasks, nodes, idx);
  const int e = it.nd.e, F = e + it.nd.b;
  const int k0 = FWD ? 0 : ((it.flags & kEntVisited) ? 0 : e);
  const int rows = it.rows;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int RP = 1;
  while (RP < rows) RP <<= 1;
  const int CG = 64 / RP;
  const int r = lane & (RP - 1), cg = lane / RP;
  const bool ok = r < rows;
  const long long ld = FWD ? it.nd.b : e;
  const int kb = k0 + w * CG + cg;
  const int K = ok ? (FWD ? e : F) : kb;
  // a lane without columns reads the factor base (valid) instead of its row
  const T2* M = K > kb ? factor + it.tk.facBase + (FWD ? it.nd.fwdOff : it.nd.bwdOff) + it.r0 + r : factor;
  const bool storer = w == 0 && cg == 0 && ok;
  T2 pre[D];
  rowLoad<T2, D>(M, ld, kb, K, W * CG, pre);
  // the children's updates of a forward row are added by the storing lanes
  T2 add[NR];
  int dof = 0;
  if constexpr (FWD)
    fwdRowAdd<T2, NR>(it, storer, r, poolMap, G, gs, add);
  else
    dof = poolE[it.nd.E + (ok ? it.r0 + r : 0)];
  if constexpr (FWD)
    fwdGather<T2, NR>(it, true, threadIdx.x, 64 * W, poolE, poolMap, X, Rb, G, xs, gs, v);
  else
    bwdGather<T2, NR>(it, true, k0 + threadIdx.x, 64 * W, poolE, poolB, poolMap, X, Xo, Rb, G, xs, gs, v);
  __syncthreads();
  T2 acc[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) acc[q] = FWD ? add[q] : czero<T2>();
  rowDotPre<T2, NR, D>(M, ld, kb, K, W * CG, v, pad, pre, acc);
  for (int off = RP; off < 64; off <<= 1) {
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = cadd(acc[q], shfl_xor2(acc[q], off));
  }
  if (cg == 0 && ok) {
#pragma unroll
    for (int q = 0; q < NR; ++q) red[(w * 64 + r) * NR + q] = acc[q];
  }
  __syncthreads();
  if (storer) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      T2 s = red[r * NR + q];
      for (int ww = 1; ww < W; ++ww) s = cadd(s, red[(ww * 64 + r) * NR + q]);
      acc[q] = s;
    }
  }
  if constexpr (FWD)
    fwdStore<T2, NR>(it, storer, r, acc, G, gs);
  else
    bwdStore<T2, NR>(it, storer, dof, acc, Xo, xs);
}

// ------------------------------------------------------------------ MFMA
// Packed entries of FP32 tasks of 8 right-hand sides on the matrix cores.
// The gather is the packed kernel's (RP >= 16 lanes per entry, P = 64/RP
// entries per wavefront); then the wavefront's 64 lanes are four 16-row
// tiles, each of one entry, and every tile takes two v_mfma_f32_16x16x4f32
// per k-step of 4 columns: with A = Ar + i Ai the tile's factor rows and V =
// Vr + i Vi the staged right-hand sides,
//   C = Ar [Vr Vi] + Ai [-Vi Vr] = [Re(A V) Im(A V)]   (16 x 16 real).
// Operand layout of the instruction: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; it holds C[4 (l / 16) + r][l % 16] in result register r.
// The entries' scalars the other lanes of a tile need are staged in LDS, and
// the products return to one row per lane through LDS for the packed
// kernel's scatter.
// Per wavefront after the staging: its entries' MfmaEnt records, then the
// 64 x 16 product (rows padded to kMfmaRow floats); kMfmaOut float2 in all.
constexpr int kMfmaRow = 20;
struct MfmaEnt;
struct MfmaEnt {
  long long mOff;  // factor offset of the entry's first row
  int ld, k0, K, rows;
};
typedef float lswF4 __attribute__((ext_vector_type(4)));
// float2 words per wavefront: 4 MfmaEnt (at most 4 entries) + 64 x kMfmaRow floats
constexpr int kMfmaOut = (4 * (int)sizeof(MfmaEnt) + 64 * kMfmaRow * 4) / 8;

// One pass of k_nd_mfma's k-loop over two tiles X and Y whose factor rows
// this wavefront loads, and, when `partners`, over tiles PX and PY of the
// same factor rows (the same node rows of another task of the cell, staged
// in another region): each loaded factor operand then feeds the matrix-core
// products of two tasks.
template <int X, int Y, int PX, int PY>
__device__ __forceinline__ void mfmaPairPass(const float2* const (&tM)[4], const int (&tk0)[4], const int (&tK)[4],
                                             const int (&tld)[4], const int (&tg)[4], const float2* vw, int cols,
                                             int pad, int li, int lk, bool partners, lswF4 (&acc)[4]) {
  constexpr int NR = 8, KS = 4;
  const bool lo = li < 8;
  const int km = max(tK[X] - tk0[X], tK[Y] - tk0[Y]);
  float2 ca[KS][2], na[KS][2];
  auto loadA = [&](int kb, float2 (&dst)[KS][2]) {
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      const int cx = tk0[X] + kb + 4 * st + lk, cy = tk0[Y] + kb + 4 * st + lk;
      dst[st][0] = tM[X][(long long)(cx < tK[X] ? cx : 0) * tld[X]];
      dst[st][1] = tM[Y][(long long)(cy < tK[Y] ? cy : 0) * tld[Y]];
    }
  };
  auto bop = [&](int t, int kb, int st, float& b1, float& b2) {
    const int col = tk0[t] + kb + 4 * st + lk;
    const float2 w = vw[((size_t)tg[t] * cols + (col < tK[t] ? col : pad)) * NR + (li & 7)];
    b1 = lo ? w.x : w.y;
    b2 = lo ? -w.y : w.x;
  };
  loadA(0, ca);
  for (int kb = 0; kb < km; kb += 4 * KS) {
    const bool more = kb + 4 * KS < km;
    if (more) loadA(kb + 4 * KS, na);
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      float b1, b2;
      bop(X, kb, st, b1, b2);
      acc[X] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][0].x, b1, acc[X], 0, 0, 0);
      acc[X] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][0].y, b2, acc[X], 0, 0, 0);
      bop(Y, kb, st, b1, b2);
      acc[Y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][1].x, b1, acc[Y], 0, 0, 0);
      acc[Y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][1].y, b2, acc[Y], 0, 0, 0);
      if (partners) {
        bop(PX, kb, st, b1, b2);
        acc[PX] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][0].x, b1, acc[PX], 0, 0, 0);
        acc[PX] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][0].y, b2, acc[PX], 0, 0, 0);
        bop(PY, kb, st, b1, b2);
        acc[PY] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][1].x, b1, acc[PY], 0, 0, 0);
        acc[PY] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][1].y, b2, acc[PY], 0, 0, 0);
      }
    }
    if (!more) break;
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      ca[st][0] = na[st][0];
      ca[st][1] = na[st][1];
    }
  }
}

// PAIR selects how the four tiles of a wavefront may share factor rows:
// 0 never (each tile loads its own); 1 (16 lanes per entry) entries 0, 1 and
// entries 2, 3 are the same node rows of two tasks of one cell; 2 (32 lanes
// per entry) entries 0 and 1 are.  The planner orders a launch's entries so
// (paired entries first); a wavefront whose entries do not pair up (checked
// here, uniformly) loads every tile itself in two passes.
template <bool FWD, int PAIR>
__global__ __launch_bounds__(256) void k_nd_mfma(const SolveTaskDev* __restrict__ tasks,
                                                 const SolveNode* __restrict__ nodes,
                                                 const SolveEntry* __restrict__ ent, int nent, int rpLog, int cols,
                                                 const int* __restrict__ poolE, const int* __restrict__ poolB,
                                                 const int* __restrict__ poolMap, const float2* __restrict__ factor,
                                                 const float2* X, float2* Xo, float2* __restrict__ Rb,
                                                 float2* __restrict__ G, long long xs, long long gs) {
  constexpr int NR = 8;
  extern __shared__ char smem_raw[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int RP = 1 << rpLog, P = 64 >> rpLog, wpb = blockDim.x >> 6;
  const int grp = lane >> rpLog, r = lane & (RP - 1);
  float2* vw = reinterpret_cast<float2*>(smem_raw) + (size_t)wave * P * NR * cols;  // the wavefront's regions
  float2* v = vw + (size_t)grp * NR * cols;
  MfmaEnt* info = reinterpret_cast<MfmaEnt*>(reinterpret_cast<float2*>(smem_raw) + (size_t)wpb * P * NR * cols +
                                              (size_t)wave * kMfmaOut);
  const int pad = cols - 1;
  if (r < NR) v[pad * NR + r] = czero<float2>();  // RP >= 16 > NR
  const long long waveId = (long long)xcdRemap(blockIdx.x, gridDim.x) * wpb + wave;
  const long long idx = waveId * P + grp;
  const bool valid = idx < nent;
  const Item it = loadItem(ent, tasks, nodes, valid ? idx : nent - 1);
  const int e = it.nd.e, F = e + it.nd.b;
  const int k0 = FWD ? 0 : ((it.flags & kEntVisited) ? 0 : e);
  // (a forward root entry of rows = 0 only keeps r_E: no product)
  const int K = valid && it.rows > 0 ? (FWD ? e : F) : k0;
  if (r == 0) {
    MfmaEnt& m = info[grp];
    // an entry without columns points at the factor base (valid) with ld 0
    m.mOff = K > k0 ? it.tk.facBase + (FWD ? it.nd.fwdOff : it.nd.bwdOff) + it.r0 : 0;
    m.ld = K > k0 ? (FWD ? it.nd.b : e) : 0;
    m.k0 = k0;
    m.K = K;
    m.rows = valid ? it.rows : 0;
  }
  const bool ok = valid && r < it.rows;
  float2 add[NR];
  int dof = 0;
  if constexpr (FWD)
    fwdRowAdd<float2, NR>(it, ok, r, poolMap, G, gs, add);
  else
    dof = poolE[it.nd.E + (ok ? it.r0 + r : 0)];
  if constexpr (FWD)
    fwdGather<float2, NR>(it, valid, r, RP, poolE, poolMap, X, Rb, G, xs, gs, v);
  else
    bwdGather<float2, NR>(it, valid, k0 + r, RP, poolE, poolB, poolMap, X, Xo, Rb, G, xs, gs, v);
  waveSync();
  // this lane's part of tile t: row rb_t + (lane & 15) of the tile's entry
  const int li = lane & 15, lk = lane >> 4;
  int tg[4], trb[4], tk0[4], tK[4], tld[4];
  const float2* tM[4];
  int kmax = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    tg[t] = (16 * t) >> rpLog;
    trb[t] = (16 * t) & (RP - 1);
    const MfmaEnt& m = info[tg[t]];
    tk0[t] = m.k0;
    tK[t] = m.K;
    tld[t] = m.ld;
    const int row = trb[t] + li;
    tM[t] = factor + m.mOff + (row < m.rows ? row : 0);
    kmax = max(kmax, m.K - m.k0);
  }
  lswF4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = lswF4{0.f, 0.f, 0.f, 0.f};
  if constexpr (PAIR == 0) {
    const bool lo = li < 8;
    // factor operands in batches of KS k-steps, the next batch loaded while
    // the current one is multiplied (the loop is uniform over the wavefront)
    constexpr int KS = 4;
    float2 ca[KS][4], na[KS][4];
    auto loadA = [&](int kb, float2 (&dst)[KS][4]) {
#pragma unroll
      for (int st = 0; st < KS; ++st)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int col = tk0[t] + kb + 4 * st + lk;
          dst[st][t] = tM[t][(long long)(col < tK[t] ? col : 0) * tld[t]];
        }
    };
    loadA(0, ca);
    for (int kb = 0; kb < kmax; kb += 4 * KS) {
      const bool more = kb + 4 * KS < kmax;
      if (more) loadA(kb + 4 * KS, na);
#pragma unroll
      for (int st = 0; st < KS; ++st)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int col = tk0[t] + kb + 4 * st + lk;
          const float2 w = vw[((size_t)tg[t] * cols + (col < tK[t] ? col : pad)) * NR + (li & 7)];
          const float b1 = lo ? w.x : w.y;
          const float b2 = lo ? -w.y : w.x;
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][t].x, b1, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][t].y, b2, acc[t], 0, 0, 0);
        }
      if (!more) break;
#pragma unroll
      for (int st = 0; st < KS; ++st)
#pragma unroll
        for (int t = 0; t < 4; ++t) ca[st][t] = na[st][t];
    }
  } else {
    constexpr int X1 = PAIR == 1 ? 2 : 1, Q0 = PAIR == 1 ? 1 : 2;
    auto sameRows = [&](int a, int b) {
      return info[a].mOff == info[b].mOff && info[a].ld == info[b].ld && info[a].k0 == info[b].k0 &&
             info[a].K == info[b].K && info[a].rows == info[b].rows;
    };
    const bool paired = PAIR == 1 ? (sameRows(0, 1) && sameRows(2, 3)) : sameRows(0, 1);
    mfmaPairPass<0, X1, Q0, 3>(tM, tk0, tK, tld, tg, vw, cols, pad, li, lk, paired, acc);
    if (!paired) mfmaPairPass<Q0, 3, 0, X1>(tM, tk0, tK, tld, tg, vw, cols, pad, li, lk, false, acc);
  }
  // back to one row per lane (lane l is row l & (RP - 1) of entry l >> rpLog,
  // which is row l % 16 of tile l / 16) through LDS, then the packed
  // kernel's scatter: children's updates and destinations were loaded with
  // the gather
  float* cb = reinterpret_cast<float*>(info + 4);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) cb[(16 * t + 4 * lk + rr) * kMfmaRow + li] = acc[t][rr];
  waveSync();
  float2 res[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    res[q] = make_float2(cb[lane * kMfmaRow + q], cb[lane * kMfmaRow + 8 + q]);
    if constexpr (FWD) res[q] = cadd(res[q], add[q]);
  }
  if constexpr (FWD)
    fwdStore<float2, NR>(it, ok, r, res, G, gs);
  else
    bwdStore<float2, NR>(it, ok, dof, res, Xo, xs);
}

// Few large entries of FP32 tasks of 8 right-hand sides on the matrix cores:
// one workgroup per entry and W wavefronts splitting its columns in
// interleaved k-steps of 4 (wavefront w takes k-steps w, w + W, w + 2W, ...).
// The right-hand side is gathered once for all of them (as k_nd_wide); each
// wavefront forms its part of the 64 x 16 product in four 16 x 16 tiles with
// k_nd_mfma's operand layout; after a barrier the staging holds the W
// partial products, which wavefront 0 sums in wavefront order (a fixed order
// of additions) before the store.  LDS: max(NR x cols staging, W x 64 x
// kMfmaRow floats).
template <bool FWD, int W>
__global__ __launch_bounds__(64 * W) void k_nd_mfma_wide(const SolveTaskDev* __restrict__ tasks,
                                                         const SolveNode* __restrict__ nodes,
                                                         const SolveEntry* __restrict__ ent, int nent, int cols,
                                                         const int* __restrict__ poolE, const int* __restrict__ poolB,
                                                         const int* __restrict__ poolMap,
                                                         const float2* __restrict__ factor, const float2* X,
                                                         float2* Xo, float2* __restrict__ Rb, float2* __restrict__ G,
                                                         long long xs, long long gs) {
  constexpr int NR = 8;
  extern __shared__ char smem_raw[];
  float2* v = reinterpret_cast<float2*>(smem_raw);
  const int pad = cols - 1;
  if (threadIdx.x < NR) v[pad * NR + threadIdx.x] = czero<float2>();  // ordered by the barrier after the gather
  const long long idx = xcdRemap(blockIdx.x, gridDim.x);
  if (idx >= nent) return;  // the whole workgroup
  const Item it = loadItem(ent, tasks, nodes, idx);
  const int e = it.nd.e, F = e + it.nd.b;
  const int k0 = FWD ? 0 : ((it.flags & kEntVisited) ? 0 : e);
  const int rows = it.rows;
  // (a forward root entry of rows = 0 only keeps r_E: no product)
  const int K = rows > 0 ? (FWD ? e : F) : k0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const long long ld = FWD ? it.nd.b : e;
  // an entry without columns reads the factor base (valid)
  const float2* M0 = factor + (K > k0 ? it.tk.facBase + (FWD ? it.nd.fwdOff : it.nd.bwdOff) + it.r0 : 0);
  const float2* tM[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int row = 16 * t + li;
    tM[t] = M0 + (row < rows ? row : 0);
  }
  const bool storer = w == 0 && lane < rows;
  float2 add[NR];
  int dof = 0;
  if constexpr (FWD)
    fwdRowAdd<float2, NR>(it, storer, lane, poolMap, G, gs, add);
  else
    dof = poolE[it.nd.E + (storer ? it.r0 + lane : 0)];
  const bool lo = li < 8;
  const int nsteps = (K - k0 + 3) / 4;
  const int nw = nsteps > w ? (nsteps - w + W - 1) / W : 0;  // this wavefront's k-steps (uniform)
  constexpr int KS = 4;
  float2 ca[KS][4], na[KS][4];
  auto colOf = [&](int s) { return k0 + 4 * (s * W + w) + lk; };
  auto loadA = [&](int s0, float2 (&dst)[KS][4]) {
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      const int col = colOf(s0 + st);
#pragma unroll
      for (int t = 0; t < 4; ++t) dst[st][t] = tM[t][(long long)(col < K ? col : 0) * ld];
    }
  };
  // the first factor operands do not depend on the gather: in flight with it
  loadA(0, ca);
  if constexpr (FWD)
    fwdGather<float2, NR>(it, true, threadIdx.x, 64 * W, poolE, poolMap, X, Rb, G, xs, gs, v);
  else
    bwdGather<float2, NR>(it, true, k0 + threadIdx.x, 64 * W, poolE, poolB, poolMap, X, Xo, Rb, G, xs, gs, v);
  __syncthreads();
  lswF4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = lswF4{0.f, 0.f, 0.f, 0.f};
  for (int s0 = 0; s0 < nw; s0 += KS) {
    const bool more = s0 + KS < nw;
    if (more) loadA(s0 + KS, na);
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      const int col = colOf(s0 + st);
      const float2 x = v[(col < K ? col : pad) * NR + (li & 7)];
      const float b1 = lo ? x.x : x.y;
      const float b2 = lo ? -x.y : x.x;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][t].x, b1, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[st][t].y, b2, acc[t], 0, 0, 0);
      }
    }
    if (!more) break;
#pragma unroll
    for (int st = 0; st < KS; ++st)
#pragma unroll
      for (int t = 0; t < 4; ++t) ca[st][t] = na[st][t];
  }
  __syncthreads();  // every wavefront is done with the staging
  float* red = reinterpret_cast<float*>(smem_raw);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) red[(w * 64 + 16 * t + 4 * lk + rr) * kMfmaRow + li] = acc[t][rr];
  __syncthreads();
  if (w != 0) return;
  float2 res[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    float re = red[lane * kMfmaRow + q], im = red[lane * kMfmaRow + 8 + q];
    for (int ww = 1; ww < W; ++ww) {
      re += red[(ww * 64 + lane) * kMfmaRow + q];
      im += red[(ww * 64 + lane) * kMfmaRow + 8 + q];
    }
    res[q] = make_float2(re, im);
    if constexpr (FWD) res[q] = cadd(res[q], add[q]);
  }
  if constexpr (FWD)
    fwdStore<float2, NR>(it, storer, lane, res, G, gs);
  else
    bwdStore<float2, NR>(it, storer, dof, res, Xo, xs);
}

// ------------------------------------------------------------------ grouped
// Matrix-core solves of FP32 tasks of 8 right-hand sides with the entries of
// one (cell, node, row chunk) of up to kGrpTasks tasks in one wavefront (a
// group): the column indices, the children's maps and the factor tiles are
// loaded once for all of them, and every factor operand feeds the products
// of every task.  The columns are walked in chunks of kGrpCols: the gather of
// chunk c+1 (index loads one chunk earlier still) and its factor operands are
// in flight while the products of chunk c run, from LDS staging that holds
// the chunk of every task (double-buffered).  The products are moved to one
// row per lane through LDS for the stores, one task at a time (a forward
// row's children's updates for the next task are loaded meanwhile).
constexpr int kGrpTasks = 4;
constexpr int kGrpCols = 16;
constexpr int kGrpWaves = 4;  // wavefronts per workgroup
// LDS per wavefront (float2 words): two staging buffers of kGrpTasks x
// kGrpCols x 8, reused after the columns for the 64 x kMfmaRow float rows
constexpr int kGrpStage = kGrpTasks * kGrpCols * 8;
constexpr int kGrpRows = 64 * kMfmaRow / 2;  // the 64 product rows of one task (float2 words)
// two staging buffers, the rows reusing them; one staging buffer and the
// rows beside it (single-chunk launches)
constexpr int kGrpLds = (2 * kGrpStage > kGrpStage + kGrpRows) ? 2 * kGrpStage : kGrpStage + kGrpRows;

// float offset of component li (0..7: real parts, 8..15: imaginary parts of
// the 8 right-hand sides) in a row of 8 interleaved complex values
__device__ __forceinline__ int grpComp(int li) { return li < 8 ? 2 * li : 2 * (li - 8) + 1; }

// ONE: every entry of the launch has at most kGrpCols columns.  Its columns
// are staged at once, and the tasks then run one after the other (products,
// rows, stores) with one task's accumulators, which leaves the registers for
// twice the wavefronts per SIMD.
template <bool FWD, bool ONE>
__global__ __launch_bounds__(64 * kGrpWaves) void k_nd_mfma_grp(
    const GroupRec* __restrict__ grp, int ngrp, const int* __restrict__ poolE, const int* __restrict__ poolB,
    const int* __restrict__ poolMap, const float2* __restrict__ factor, const float2* X, float2* Xo,
    float2* __restrict__ Rb, float2* __restrict__ G, long long xs, long long gs) {
  constexpr int NR = 8, TM = kGrpTasks, KC = kGrpCols;
  __shared__ float2 lds[kGrpWaves][kGrpLds];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const long long gid = (long long)xcdRemap(blockIdx.x, gridDim.x) * kGrpWaves + wave;
  if (gid >= ngrp) return;  // this wavefront only (no workgroup barriers below)
  float2* st = lds[wave];
  // everything about the group in one record
  const GroupRec& gr = grp[gid];
  const SolveNode nd = gr.nd;
  const int T = gr.ntasks;
  const int r0 = gr.r0rows & 0xffff, rows = gr.r0rows >> 16;
  const int e = nd.e, F = e + nd.b;
  int tTask[TM], tFlags[TM];
  int anyVis = 0;
#pragma unroll
  for (int t = 0; t < TM; ++t) {
    tTask[t] = gr.task[t];
    tFlags[t] = gr.flags[t];
    anyVis |= tFlags[t] & (kEntVisited | kEntRootRhs);
  }
  const long long facBase = gr.facBase;  // one cell
  // columns: forward r_E (e of them); backward [r_E; x_B], from e on when no
  // task has r_E there
  const int kb = FWD ? 0 : (anyVis ? 0 : e);
  const int ke = FWD ? e : F;
  const int nt = (rows + 15) >> 4;  // 16-row tiles with rows
  const int li = lane & 15, lk = lane >> 4;
  const bool lo = li < 8;
  const long long ld = FWD ? nd.b : e;
  const float2* M0 = factor + facBase + (FWD ? nd.fwdOff : nd.bwdOff) + r0;
  const float2* tM[4];
#pragma unroll
  for (int tl = 0; tl < 4; ++tl) {
    const int row = 16 * tl + li;
    tM[tl] = M0 + (row < rows ? row : 0);
  }
  // this lane's gather pair of every chunk: column kk of task tg
  const int kk = lane & (KC - 1), tg = lane / KC;
  const bool tgOk = tg < T;
  const int myTask = tg == 0 ? tTask[0] : tg == 1 ? tTask[1] : tg == 2 ? tTask[2] : tTask[3];
  const int myFlags = tg == 0 ? tFlags[0] : tg == 1 ? tFlags[1] : tg == 2 ? tFlags[2] : tFlags[3];
  const bool myRoot = !FWD && (myFlags & kEntRootRhs);
  const bool myVis = FWD || (myFlags & (kEntVisited | kEntRootRhs));
  const bool myC1 = (myFlags & kEntChild1) != 0, myC2 = (myFlags & kEntChild2) != 0;
  const bool keepR = FWD && r0 == 0;
  // children's maps are read for a forward gather, or a backward one of a
  // root forming its r_E (uniform)
  const bool kidsG = waveAny(tgOk && (FWD || myRoot) && (myC1 || myC2));
  const int* Ep = poolE + nd.E;
  const int* Bp = poolB + nd.B;
  const int* m1 = poolMap + nd.map;

  struct Idx {
    int dof, a1, a2;
    bool act;
  };
  auto loadIdx = [&](int kbase) __attribute__((always_inline)) {
    Idx ix;
    const int k = kbase + kk;
    ix.act = tgOk && k < ke;
    const int kc = k < ke ? k : kb;  // a valid column (kb < ke inside the loop)
    ix.dof = kc < e ? Ep[kc] : Bp[kc - e];
    ix.a1 = -1;
    ix.a2 = -1;
    if (kidsG) {
      const int km = kc < e ? kc : 0;
      ix.a1 = m1[km];
      ix.a2 = m1[F + km];
      if (kc >= e) ix.a1 = ix.a2 = -1;
    }
    // a pair that is past the columns, or a backward r_E column of a task
    // without r_E, stages zero
    if (!FWD && kc < e && !myVis) ix.act = false;
    return ix;
  };
  struct Val {
    float2 x[NR], g1[NR], g2[NR];
    bool h1, h2, r;
  };
  auto loadVal = [&](const Idx& ix, int kbase, Val& v) __attribute__((always_inline)) {
    const int k = kbase + kk;
    const bool isR = k < e;  // an r_E column (past ke: any, masked)
    // forward, or a backward root: X; backward r_E: the kept R; x_B: Xo
    // (selected as integers: a select of pointers went through scratch)
    unsigned long long sa = (unsigned long long)X;
    if (!FWD) {
      sa = (isR && !myRoot) ? (unsigned long long)Rb : sa;
      sa = !isR ? (unsigned long long)Xo : sa;
    }
    const float2* src = reinterpret_cast<const float2*>(sa);
    loadNR<float2, NR>(src + ((long long)myTask * xs + ix.dof) * NR, v.x);
    v.h1 = (FWD || myRoot) && isR && myC1 && ix.a1 >= 0;
    v.h2 = (FWD || myRoot) && isR && myC2 && ix.a2 >= 0;
    v.r = ix.act;
    if (kidsG) {
      loadNR<float2, NR>(G + ((long long)myTask * gs + (v.h1 ? nd.g1 + ix.a1 : 0)) * NR, v.g1);
      loadNR<float2, NR>(G + ((long long)myTask * gs + (v.h2 ? nd.g2 + ix.a2 : 0)) * NR, v.g2);
    }
  };
  auto putVal = [&](const Idx& ix, const Val& v, float2* buf) __attribute__((always_inline)) {
    float2 o[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      float2 w = v.x[q];
      if (kidsG) w = csub(csub(w, cmask(v.g1[q], v.h1)), cmask(v.g2[q], v.h2));
      o[q] = cmask(w, v.r);
    }
    // every lane writes its pair (a task past T writes its own, unread region)
#pragma unroll
    for (int q = 0; q < NR; ++q) buf[(tg * KC + kk) * NR + q] = o[q];
    if (keepR && v.r) storeNR<float2, NR>(Rb + ((long long)myTask * xs + ix.dof) * NR, o);
  };
  auto loadA = [&](int kbase, float2 (&a)[4][4]) __attribute__((always_inline)) {
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const int col = kbase + 4 * s4 + lk;
      const long long cc = col < ke ? col : kb;
#pragma unroll
      for (int tl = 0; tl < 4; ++tl)
        if (tl < nt) a[s4][tl] = tM[tl][cc * ld];  // (nt is uniform)
    }
  };

  // the stores: one row per lane (row `lane` of the chunk); their
  // destinations (backward) or the positions of the children's updates
  // they add (forward) are loaded now, while the products run
  const bool ok = lane < rows;
  float* cb = reinterpret_cast<float*>(st);
  int dof = 0, ra1 = -1, ra2 = -1;
  if (nt > 0) {
    if constexpr (FWD) {
      const int f = e + r0 + (ok ? lane : 0);
      ra1 = m1[f];
      ra2 = m1[F + f];
    } else {
      dof = Ep[r0 + (ok ? lane : 0)];
    }
  }
  auto loadAdd = [&](int t, float2 (&c1)[NR], float2 (&c2)[NR], bool& h1, bool& h2) __attribute__((always_inline)) {
    const int fl = tFlags[t];
    h1 = ok && (fl & kEntChild1) && ra1 >= 0;
    h2 = ok && (fl & kEntChild2) && ra2 >= 0;
    loadNR<float2, NR>(G + ((long long)tTask[t] * gs + (h1 ? nd.g1 + ra1 : 0)) * NR, c1);
    loadNR<float2, NR>(G + ((long long)tTask[t] * gs + (h2 ? nd.g2 + ra2 : 0)) * NR, c2);
  };
  if constexpr (ONE) {
    if (nt == 0 && kb >= ke) return;
    float2 a[4][4];
    if (kb < ke) {
      const Idx ix = loadIdx(kb);
      Val v;
      loadVal(ix, kb, v);
      loadA(kb, a);
      putVal(ix, v, st);
    }
    if (nt == 0) return;  // a forward root entry only keeps r_E
    float* rws = reinterpret_cast<float*>(st + kGrpStage);
    float2 c1C[NR], c2C[NR];
    bool h1C = false, h2C = false;
    if constexpr (FWD) loadAdd(0, c1C, c2C, h1C, h2C);
    waveSync();
#pragma unroll
    for (int t = 0; t < TM; ++t) {
      if (t >= T) break;
      float2 c1N[NR], c2N[NR];
      bool h1N = false, h2N = false;
      if constexpr (FWD) {
        if (t + 1 < T) loadAdd(t + 1, c1N, c2N, h1N, h2N);
      }
      lswF4 acc1[4];
#pragma unroll
      for (int tl = 0; tl < 4; ++tl) acc1[tl] = lswF4{0.f, 0.f, 0.f, 0.f};
      if (kb < ke) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const float2 w = st[(t * KC + 4 * s4 + lk) * NR + (li & 7)];
          const float b1 = lo ? w.x : w.y;
          const float b2 = lo ? -w.y : w.x;
#pragma unroll
          for (int tl = 0; tl < 4; ++tl) {
            if (tl >= nt) break;
            acc1[tl] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4][tl].x, b1, acc1[tl], 0, 0, 0);
            acc1[tl] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4][tl].y, b2, acc1[tl], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int tl = 0; tl < 4; ++tl)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) rws[(16 * tl + 4 * lk + rr) * kMfmaRow + li] = acc1[tl][rr];
      waveSync();
      float2 res[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        res[q] = make_float2(rws[lane * kMfmaRow + q], rws[lane * kMfmaRow + 8 + q]);
        if constexpr (FWD) res[q] = cadd(res[q], cadd(cmask(c1C[q], h1C), cmask(c2C[q], h2C)));
      }
      waveSync();  // the rows are read before the next task's are written
      if constexpr (FWD) {
        float2* dst = ok ? G + ((long long)tTask[t] * gs + nd.gOff + r0 + lane) * NR : trashRow<float2, NR>();
        storeNR<float2, NR>(dst, res);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          c1C[q] = c1N[q];
          c2C[q] = c2N[q];
        }
        h1C = h1N;
        h2C = h2N;
      } else {
        storeNR<float2, NR>(ok ? Xo + ((long long)tTask[t] * xs + dof) * NR : trashRow<float2, NR>(), res);
      }
    }
    return;
  }

  lswF4 acc[TM][4];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int tl = 0; tl < 4; ++tl) acc[t][tl] = lswF4{0.f, 0.f, 0.f, 0.f};

  if (kb < ke) {
    Idx ixC = loadIdx(kb);
    Val vC;
    loadVal(ixC, kb, vC);
    Idx ixN = loadIdx(kb + KC);
    float2 aC[4][4], aN[4][4];
    loadA(kb, aC);
    int bi = 0;
    for (int kbase = kb;; kbase += KC) {
      const bool more = kbase + KC < ke;
      putVal(ixC, vC, st + bi * kGrpStage);
      if (more) {
        loadVal(ixN, kbase + KC, vC);
        ixC = ixN;
        ixN = loadIdx(kbase + 2 * KC);
        loadA(kbase + KC, aN);
      }
      waveSync();
      const float2* buf = st + bi * kGrpStage;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
#pragma unroll
        for (int t = 0; t < TM; ++t) {
          if (t >= T) break;
          const float2 w = buf[(t * KC + 4 * s4 + lk) * NR + (li & 7)];
          const float b1 = lo ? w.x : w.y;
          const float b2 = lo ? -w.y : w.x;
#pragma unroll
          for (int tl = 0; tl < 4; ++tl) {
            if (tl >= nt) break;
            acc[t][tl] = __builtin_amdgcn_mfma_f32_16x16x4f32(aC[s4][tl].x, b1, acc[t][tl], 0, 0, 0);
            acc[t][tl] = __builtin_amdgcn_mfma_f32_16x16x4f32(aC[s4][tl].y, b2, acc[t][tl], 0, 0, 0);
          }
        }
      }
      if (!more) break;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
        for (int tl = 0; tl < 4; ++tl) aC[s4][tl] = aN[s4][tl];
      bi ^= 1;
    }
  }
  if (nt == 0) return;  // a forward root entry only keeps r_E
  waveSync();  // the staging is reused for the rows below
  float2 c1C[NR], c2C[NR];
  bool h1C = false, h2C = false;
  if constexpr (FWD) loadAdd(0, c1C, c2C, h1C, h2C);
#pragma unroll
  for (int t = 0; t < TM; ++t) {
    if (t >= T) break;
    float2 c1N[NR], c2N[NR];
    bool h1N = false, h2N = false;
    if constexpr (FWD) {
      if (t + 1 < T) loadAdd(t + 1, c1N, c2N, h1N, h2N);
    }
#pragma unroll
    for (int tl = 0; tl < 4; ++tl)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) cb[(16 * tl + 4 * lk + rr) * kMfmaRow + li] = acc[t][tl][rr];
    waveSync();
    float2 res[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      res[q] = make_float2(cb[lane * kMfmaRow + q], cb[lane * kMfmaRow + 8 + q]);
      if constexpr (FWD) res[q] = cadd(res[q], cadd(cmask(c1C[q], h1C), cmask(c2C[q], h2C)));
    }
    waveSync();  // the rows are read before the next task's are written
    if constexpr (FWD) {
      float2* dst = ok ? G + ((long long)tTask[t] * gs + nd.gOff + r0 + lane) * NR : trashRow<float2, NR>();
      storeNR<float2, NR>(dst, res);
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        c1C[q] = c1N[q];
        c2C[q] = c2N[q];
      }
      h1C = h1N;
      h2C = h2N;
    } else {
      storeNR<float2, NR>(ok ? Xo + ((long long)tTask[t] * xs + dof) * NR : trashRow<float2, NR>(), res);
    }
  }
}
}  // namespace

// ------------------------------------------------------------------ plans
SolvePlan::~SolvePlan() {
  if (dBlob) (void)devFree(dBlob);
}

namespace {
bool meets(const NdNode& nd, const std::vector<Rect>& sup) {
  if (sup.empty()) return true;
  for (const Rect& r : sup)
    if (r.i0 < nd.i1 && nd.i0 < r.i1 && r.j0 < nd.j1 && nd.j0 < r.j1) return true;
  return false;
}

int ceilLog2(int x) {
  int l = 0;
  while ((1 << l) < x) ++l;
  return l;
}

}  // namespace

// Packed launches of fewer wavefronts are latency-bound and take the
// column-split shape (LSW_LATENCY_WAVES overrides the threshold; 0 turns the
// rule off); packed launches of FP32 tasks of 8 right-hand sides run on the
// matrix cores (k_nd_mfma) unless LSW_SOLVE_MFMA=0.
PlanShape PlanShape::defaults() {
  static const PlanShape d = [] {
    PlanShape p;
    if (const char* e = std::getenv("LSW_LATENCY_WAVES")) p.latencyWaves = std::max(0LL, std::atoll(e));
    if (const char* e = std::getenv("LSW_SOLVE_MFMA")) p.mfma = std::atoi(e) != 0;
    if (const char* e = std::getenv("LSW_MFMA_WIDE_COLS")) p.mfmaWideCols = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("LSW_MFMA_PAIR")) p.mfmaPair = std::atoi(e) != 0;
    if (const char* e = std::getenv("LSW_MFMA_GROUP")) p.mfmaGroup = std::atoi(e) != 0;
    if (const char* e = std::getenv("LSW_GRP_ROWS_FWD")) p.grpMinRowsFwd = std::atoi(e);
    if (const char* e = std::getenv("LSW_GRP_ROWS_BWD")) p.grpMinRowsBwd = std::atoi(e);
    if (const char* e = std::getenv("LSW_GRP_LATENCY")) p.grpLatency = std::atoi(e) != 0;
    if (const char* e = std::getenv("LSW_LATENCY_ROWS")) p.latencyRows = std::atoi(e);
    if (const char* e = std::getenv("LSW_WIDE_BUDGET")) p.wideBudget = std::max(0LL, std::atoll(e));
    if (const char* e = std::getenv("LSW_WIDE_CAP")) p.wideCap = std::max(1, std::atoi(e));
    return p;
  }();
  return d;
}

void FactorSet::buildPlan(const std::vector<SolveTask>& tasks, const std::vector<std::vector<Rect>>& supports,
                          SolvePlan& plan, const std::vector<std::vector<Rect>>* bwdSupports, bool outOfPlace,
                          const PlanShape* shapeIn) const {
  PlanShape shp = shapeIn ? *shapeIn : PlanShape::defaults();
  static const bool ptime = std::getenv("LSW_PLAN_TIME") != nullptr;
  auto clk = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
  const double pt0 = ptime ? clk() : 0;
  double pt1 = 0, pt2 = 0, pt3 = 0;
  LSW_REQUIRE(!mTrees.empty(), LSW_ERR_STATE, "FactorSet: setup() not called");
  LSW_REQUIRE(supports.empty() || supports.size() == tasks.size(), LSW_ERR_ARG, "one support per task");
  LSW_REQUIRE(!bwdSupports || bwdSupports->size() == tasks.size(), LSW_ERR_ARG, "one backward support per task");
  const int nt = (int)tasks.size();
  LSW_REQUIRE(nt > 0, LSW_ERR_ARG, "empty plan");
  plan.ntasks = nt;
  plan.hTasks = tasks;
  plan.maxNrhs = 1;
  plan.outOfPlace = outOfPlace;
  const double S = mPrecision == LSW_FP32 ? 8.0 : 16.0;
  std::vector<SolveTaskDev> tdev(nt);
  // entries per height
  std::vector<std::vector<SolveEntry>> fw(mMaxHeight + 1), bw(mMaxHeight + 1);
  std::vector<int> fcols(mMaxHeight + 1, 1), bcols(mMaxHeight + 1, 1);
  std::vector<int> frows(mMaxHeight + 1, 0), brows(mMaxHeight + 1, 0);
  std::vector<double> fbytes(mMaxHeight + 1, 0.0), bbytes(mMaxHeight + 1, 0.0);
  std::vector<double> fwork(mMaxHeight + 1, 0.0), bwork(mMaxHeight + 1, 0.0);
  std::vector<double> fflops(mMaxHeight + 1, 0.0), bflops(mMaxHeight + 1, 0.0);
  std::vector<char> on, need;
  // Algorithmic bytes: the right-hand-side and update vectors of every task,
  // and each factor block once per launch however many tasks of its cell
  // read it (they run as neighbouring entries, and the block comes from HBM
  // once): per (cell, node), the widest factor read counted so far.
  // (flat: cell c's nodes from cellNode[c] on)
  std::vector<long long> cellNode(mCells.size() + 1, 0);
  for (size_t c = 0; c < mCells.size(); ++c)
    cellNode[c + 1] = cellNode[c] + (long long)mTrees[mCells[c].shape]->nodes.size();
  std::vector<std::pair<double, double>> facSeen((size_t)cellNode.back(), {0.0, 0.0});
  auto facBytes = [&](int cell, int node, bool fwd, double bytes) {
    auto& seen = facSeen[(size_t)(cellNode[cell] + node)];
    double& c = fwd ? seen.first : seen.second;
    const double add = std::max(0.0, bytes - c);
    c = std::max(c, bytes);
    return add;
  };
  // Entries are generated cell by cell (a run of tasks of one cell), node by
  // node, row chunk by row chunk, task by task: the order the launches want
  // (the entries of one node and row chunk for all tasks of a cell adjacent),
  // so the lists need no sort when the tasks come cell by cell.
  std::vector<int> gFlags;
  for (int t0 = 0; t0 < nt;) {
    int t1 = t0 + 1;
    while (t1 < nt && tasks[t1].cell == tasks[t0].cell) ++t1;
    const int ng = t1 - t0;
    const int cell = tasks[t0].cell;
    LSW_REQUIRE(cell >= 0 && cell < (int)mCells.size(), LSW_ERR_ARG, "bad task cell");
    const int shape = mCells[cell].shape;
    const NdTree& tr = *mTrees[shape];
    const int nn = (int)tr.nodes.size();
    const int base = mShapeNodeBase[shape];
    on.assign((size_t)ng * nn, 0);
    need.assign((size_t)ng * nn, 1);
    gFlags.assign((size_t)ng, 0);
    for (int t = t0; t < t1; ++t) {
      const SolveTask& tk = tasks[t];
      LSW_REQUIRE(tk.nrhs >= 1 && tk.nrhs <= kMaxRhs, LSW_ERR_ARG, "bad task right-hand-side count");
      plan.maxNrhs = std::max(plan.maxNrhs, tk.nrhs);
      SolveTaskDev& td = tdev[t];
      std::memset(&td, 0, sizeof(td));
      td.facBase = mCells[cell].factorOff;
      td.nrhs = tk.nrhs;
      for (int q = 0; q < kMaxRhs; ++q) td.slot[q] = q < tk.nrhs ? tk.slot[q] : 0;
      char* onT = on.data() + (size_t)(t - t0) * nn;
      char* needT = need.data() + (size_t)(t - t0) * nn;
      if (supports.empty() || supports[t].empty()) {
        std::fill(onT, onT + nn, 1);
      } else {
        for (int p = 0; p < nn; ++p) onT[p] = meets(tr.nodes[p], supports[t]) ? 1 : 0;
      }
      // nodes whose solution values are needed
      if (bwdSupports) {
        const std::vector<Rect>& bs = (*bwdSupports)[t];
        for (int p = 0; p < nn; ++p) needT[p] = !bs.empty() && meets(tr.nodes[p], bs) ? 1 : 0;
      }
    }
    for (int p = 0; p < nn; ++p) {
      const NdNode& nd = tr.nodes[p];
      const int h = nd.height;
      const double e = nd.e, b = nd.b;
      bool anyFwd = false, anyBwd = false;
      for (int g = 0; g < ng; ++g) {
        const char* onT = on.data() + (size_t)g * nn;
        int flags = 0;
        if (nd.c1 >= 0 && onT[nd.c1]) flags |= kEntChild1;
        if (nd.c2 >= 0 && onT[nd.c2]) flags |= kEntChild2;
        if (onT[p]) flags |= kEntVisited;
        // A root (no boundary rows) has no forward product, only r_E to
        // form.  Its backward entries form it themselves when the solution
        // goes to another buffer, or when the root is one row chunk;
        // otherwise (in place, several chunks) a forward entry stores r_E,
        // because the chunks of its backward entries overwrite X[E] while
        // the others read it.
        if (nd.b == 0 && (outOfPlace || nd.e <= kSolveRows)) flags |= kEntRootRhs;
        gFlags[g] = flags;
        anyFwd |= onT[p] && !(flags & kEntRootRhs);
        anyBwd |= need[(size_t)g * nn + p] != 0;
      }
      if (anyFwd) {
        if (nd.b == 0) {
          for (int g = 0; g < ng; ++g)
            if (on[(size_t)g * nn + p] && !(gFlags[g] & kEntRootRhs)) fw[h].push_back({t0 + g, base + p, 0, gFlags[g]});
        } else {
          for (int r0 = 0; r0 < nd.b; r0 += kSolveRows) {
            const int rows = std::min(kSolveRows, nd.b - r0);
            for (int g = 0; g < ng; ++g)
              if (on[(size_t)g * nn + p] && !(gFlags[g] & kEntRootRhs))
                fw[h].push_back({t0 + g, base + p, r0 | (rows << 16), gFlags[g]});
            frows[h] = std::max(frows[h], rows);
          }
        }
        fcols[h] = std::max(fcols[h], nd.e);
        for (int g = 0; g < ng; ++g) {
          if (!on[(size_t)g * nn + p] || (gFlags[g] & kEntRootRhs)) continue;
          const double r = tasks[t0 + g].nrhs;
          fbytes[h] += facBytes(cell, p, true, S * b * e) + S * r * (2 * e + 2 * b);
          fwork[h] += b * e + e;
          fflops[h] += 8.0 * r * b * e;
        }
      }
      if (!anyBwd) continue;
      for (int r0 = 0; r0 < nd.e; r0 += kSolveRows) {
        const int rows = std::min(kSolveRows, nd.e - r0);
        for (int g = 0; g < ng; ++g)
          if (need[(size_t)g * nn + p]) bw[h].push_back({t0 + g, base + p, r0 | (rows << 16), gFlags[g]});
        brows[h] = std::max(brows[h], rows);
      }
      bcols[h] = std::max(bcols[h], nd.e + nd.b);
      for (int g = 0; g < ng; ++g) {
        if (!need[(size_t)g * nn + p]) continue;
        const double r = tasks[t0 + g].nrhs;
        const double kcols = on[(size_t)g * nn + p] ? e + b : b;
        bbytes[h] += facBytes(cell, p, false, S * e * kcols) + S * r * (kcols + e);
        bwork[h] += e * kcols;
        bflops[h] += 8.0 * r * e * kcols;
      }
    }
    t0 = t1;
  }
  if (ptime) pt1 = clk();
  plan.width = taskWidth(plan.maxNrhs);
  const double nrk = plan.width;
  // 16-wavefront splits exist for FP32 tasks of up to 4 right-hand sides
  shp.wideCap = std::min(shp.wideCap, (mPrecision == LSW_FP32 && plan.width <= kTaskRhs4) ? 16 : 8);
  // Within a height, the entries of one node (and row chunk) for all tasks of
  // a cell are adjacent: they read the same factor block, and adjacent
  // entries run together on one XCD (xcdRemap), so the block comes from HBM
  // once and from L2 for the other tasks.
  auto byNode = [&](const SolveEntry& a, const SolveEntry& b) {
    const int ca = tasks[a.task].cell, cb = tasks[b.task].cell;
    if (ca != cb) return ca < cb;
    if (a.node != b.node) return a.node < b.node;
    return (a.r0rows & 0xffff) < (b.r0rows & 0xffff);
  };
  // (already in this order when the tasks come cell by cell in ascending
  // cells; otherwise sorted here)
  for (auto& v : fw)
    if (!std::is_sorted(v.begin(), v.end(), byNode)) std::stable_sort(v.begin(), v.end(), byNode);
  for (auto& v : bw)
    if (!std::is_sorted(v.begin(), v.end(), byNode)) std::stable_sort(v.begin(), v.end(), byNode);
  if (ptime) pt2 = clk();
  std::vector<SolveEntry> all;
  std::vector<GroupRec> allGrp;
  {
    // room for every entry and every group a launch could form (no
    // reallocation of the large lists while they fill)
    size_t nEnt = 0, nGrp = 0;
    for (const auto* lists : {&fw, &bw})
      for (const auto& v : *lists) {
        nEnt += v.size();
        for (size_t i = 0; i < v.size();) {
          size_t j = i + 1;
          while (j < v.size() && j - i < (size_t)kGrpTasks && !byNode(v[i], v[j]) && v[i].r0rows == v[j].r0rows) ++j;
          ++nGrp;
          i = j;
        }
      }
    all.reserve(nEnt);
    allGrp.reserve(nGrp);
  }
  auto addLaunch = [&](std::vector<SolvePlan::Launch>& out, const std::vector<SolveEntry>& v, int cols, int maxRows,
                       double bytes, double work, double flops, bool fwdLaunch) {
    if (v.empty()) return;
    SolvePlan::Launch L;
    L.off = (long long)all.size();
    L.n = (int)v.size();
    L.flops = flops;
    // staging columns: the front's widest r_E / [r_E; x_B] plus a zero pad
    const int stage = cols + 1;
    L.cols = stage;
    L.bytes = bytes;
    const double per = work / v.size();
    auto ldsWide = [&](int w) { return S * nrk * ((double)stage + 64.0 * w); };
    // packed shape: RP lanes per entry (>= 8 so that the LDS stays bounded)
    // (a launch of roots only, maxRows == 0, just gathers r_E: all 64 lanes)
    int rp = maxRows == 0 ? 64 : 1 << std::max(3, ceilLog2(std::max(1, maxRows)));
    auto ldsWave = [&](int rpv) { return S * nrk * stage * (64.0 / rpv); };
    while (rp < 64 && ldsWave(rp) > kMaxLds) rp *= 2;
    const long long packedWaves = ((long long)v.size() * rp + 63) / 64;
    int W = 0;
    const bool latency = packedWaves < shp.latencyWaves && maxRows > 0 && maxRows <= shp.latencyRows;
    if (latency) {
      // Few wavefronts of short rows: the launch is a chain of memory round
      // trips, and a packed lane walks its whole row while most lanes of an
      // entry's wavefront share its few rows.  Split the columns instead (W
      // wavefronts x 64/RP column groups per entry) until a lane holds at
      // most one prefetched batch of them.  (Measured on 8- and 16-row
      // heights of a sweep step: 14-20 us -> 8-9 us; entries of 32-64 rows
      // gained nothing or lost, as the split only added wavefronts.)
      const int RPw = 1 << ceilLog2(maxRows);
      const int CG = 64 / RPw;
      W = 1;
      const int depth = kDotDepth;
      while (W < shp.wideCap && (cols + CG * W - 1) / (CG * W) > depth && (long long)v.size() * W * 2 <= 8192)
        W *= 2;
    } else {
      // few large entries: split their columns over W wavefronts
      int w = 1;
      while (w < shp.wideCap && (double)v.size() * w * 2 <= (double)shp.wideBudget && per >= 1024.0 * w * 2) w *= 2;
      // Long rows of FP32 8-right-hand-side tasks: a packed matrix-core
      // wavefront walks a 64-row entry's 200-400 columns alone, while two
      // wavefronts on halves of them keep more loads in flight (measured on
      // the 16-source lens block: 2.9-3.6 TB/s packed, 5.3-6.4 TB/s split;
      // at 127 columns packed was faster, 5.0 against 4.2 TB/s).
      if (w == 1 && S == 8.0 && nrk == kMaxRhs && shp.mfma && maxRows > 32 && cols >= shp.mfmaWideCols)
        w = std::min(2, shp.wideCap);
      if (w > 1) W = w;
    }
    while (W > 1 && ldsWide(W) > kMaxLds) W /= 2;
    // FP32 tasks of 8 right-hand sides: packed entries on the matrix cores
    const int rpm = std::max(16, rp);
    auto ldsMfma = [&](int wpb) { return wpb * ((64.0 / rpm) * S * nrk * stage + 8.0 * kMfmaOut); };
    auto ldsMfmaWide = [&](int w) { return std::max(S * nrk * stage, 4.0 * w * 64 * kMfmaRow); };
    const std::vector<SolveEntry>* ev = &v;
    std::vector<SolveEntry> paired;
    if ((W == 0 || (latency && shp.grpLatency)) && S == 8.0 && nrk == kMaxRhs && shp.mfma && shp.mfmaGroup &&
        maxRows > (fwdLaunch ? shp.grpMinRowsFwd : shp.grpMinRowsBwd)) {
      // groups: the entries of one (cell, node, row chunk), kGrpTasks at most
      L.mode = 4;
      L.goff = (long long)allGrp.size();
      size_t i = 0;
      while (i < v.size()) {
        size_t j = i + 1;
        while (j < v.size() && j - i < (size_t)kGrpTasks && !byNode(v[i], v[j]) && v[i].r0rows == v[j].r0rows) ++j;
        GroupRec g;
        std::memset(&g, 0, sizeof(g));
        g.nd = mSolveNodes[v[i].node];
        g.facBase = tdev[v[i].task].facBase;
        g.r0rows = v[i].r0rows;
        g.ntasks = (int)(j - i);
        for (int t = 0; t < kGrpTasks; ++t) {
          const SolveEntry& en = v[i + std::min<size_t>(t, j - i - 1)];
          g.task[t] = en.task;
          g.flags[t] = t < g.ntasks ? en.flags : 0;
        }
        allGrp.push_back(g);
        i = j;
      }
      L.ngroups = (int)(allGrp.size() - L.goff);
      L.one = cols <= kGrpCols;
    } else if (W == 0 && S == 8.0 && nrk == kMaxRhs && shp.mfma && ldsMfma(1) <= kMaxLds) {
      L.mode = 2;
      L.rp = rpm;
      // Entries of two tasks of one cell for the same node rows read the same
      // factor rows: put such pairs side by side in one wavefront, which
      // then loads each factor operand once for both tasks (entries of up
      // to 32 rows, when most of them pair up).  Measured on the 16-source
      // lens block: those forward heights 9% faster, backward 1.5%.  Cutting
      // 64-row entries into paired 32-row halves gained nothing forward and
      // lost backward (the halves gather the right-hand sides twice), so
      // they stay whole.
      auto ldsMfmaRp = [&](int rpv, int wpb) { return wpb * ((64.0 / rpv) * S * nrk * stage + 8.0 * kMfmaOut); };
      if (shp.mfmaPair && rpm <= 32) {
        std::vector<SolveEntry> singles;
        size_t i = 0;
        while (i < v.size()) {
          size_t j = i + 1;
          while (j < v.size() && !byNode(v[i], v[j]) && v[i].r0rows == v[j].r0rows) ++j;
          size_t k = i;
          for (; k + 1 < j; k += 2) {
            paired.push_back(v[k]);
            paired.push_back(v[k + 1]);
          }
          if (k < j) singles.push_back(v[k]);
          i = j;
        }
        const size_t npaired = paired.size();
        paired.insert(paired.end(), singles.begin(), singles.end());
        if (2 * npaired >= paired.size()) {
          L.pair = rpm == 16 ? 1 : 2;
          ev = &paired;
        }
      }
      int wpb = 4;
      while (wpb > 1 && ldsMfmaRp(L.rp, wpb) > kMaxLds) wpb /= 2;
      L.wpb = wpb;
    } else if (W > 0 && S == 8.0 && nrk == kMaxRhs && shp.mfma && ldsMfmaWide(W) <= kMaxLds) {
      // few large entries of FP32 tasks of 8 right-hand sides: their columns
      // split over W wavefronts on the matrix cores
      L.mode = 3;
      L.waves = W;
    } else if (W > 0 && ldsWide(W) <= kMaxLds) {
      L.mode = 1;
      L.waves = W;
    } else {
      L.mode = 0;
      LSW_REQUIRE(ldsWave(rp) <= kMaxLds, LSW_ERR_CAPACITY, "front too wide for the solve kernel's LDS staging");
      L.rp = rp;
      int wpb = 4;
      while (wpb > 1 && wpb * ldsWave(rp) > kMaxLds) wpb /= 2;
      L.wpb = wpb;
      // entries per lane group: keep at least packedWaveTarget() wavefronts
      const long long groups = (long long)v.size() * rp / 64 + 1;
      L.rounds = (int)std::min<long long>(16, std::max<long long>(1, groups / packedWaveTarget()));
    }
    // (a grouped launch reads its group records, not the entries)
    if (L.mode != 4) all.insert(all.end(), ev->begin(), ev->end());
    out.push_back(L);
  };
  plan.fwd.clear();
  plan.bwd.clear();
  plan.fwdBytes = 0;
  plan.bwdBytes = 0;
  plan.fwdFlops = 0;
  plan.bwdFlops = 0;
  // The forward pass only feeds the backward pass: with no solution value
  // read (sweep solves whose traces nobody uses), there is nothing to do.
  bool anyBwd = false;
  for (const auto& v : bw) anyBwd = anyBwd || !v.empty();
  if (!anyBwd) {
    for (auto& v : fw) v.clear();
    std::fill(fbytes.begin(), fbytes.end(), 0.0);
    std::fill(fflops.begin(), fflops.end(), 0.0);
  }
  for (int h = 0; h <= mMaxHeight; ++h) {
    addLaunch(plan.fwd, fw[h], fcols[h], frows[h], fbytes[h], fwork[h], fflops[h], true);
    plan.fwdBytes += fbytes[h];
    if (!fw[h].empty()) plan.fwdFlops += fflops[h];
  }
  for (int h = mMaxHeight; h >= 0; --h) {
    addLaunch(plan.bwd, bw[h], bcols[h], brows[h], bbytes[h], bwork[h], bflops[h], false);
    plan.bwdBytes += bbytes[h];
    if (!bw[h].empty()) plan.bwdFlops += bflops[h];
  }
  if (std::getenv("LSW_PLAN_LOG")) {
    for (int pass = 0; pass < 2; ++pass)
      for (const auto& L : pass ? plan.bwd : plan.fwd)
        std::fprintf(stderr, "plan %s tasks %d nrhs %d: entries %d mode %d rp %d rounds %d wpb %d waves %d cols %d "
                     "bytes %.0f\n", pass ? "bwd" : "fwd", nt, plan.maxNrhs, L.n, L.mode, L.rp, L.rounds, L.wpb,
                     L.waves, L.cols, L.bytes);
  }
  if (ptime) pt3 = clk();
  if (plan.dBlob) (void)devFree(plan.dBlob);
  plan.dBlob = nullptr;
  plan.dTasks = nullptr;
  plan.dEntries = nullptr;
  plan.dSlotTQ = nullptr;
  int maxSlot = 0;
  for (const auto& tk : tasks)
    for (int q = 0; q < tk.nrhs; ++q) maxSlot = std::max(maxSlot, tk.slot[q]);
  plan.slotTQ.assign((size_t)maxSlot + 1, -1);
  for (int t = 0; t < nt; ++t)
    for (int q = 0; q < tasks[t].nrhs; ++q) plan.slotTQ[tasks[t].slot[q]] = t * plan.width + q;
  // tasks, entries and the slot map in one allocation and one upload (a
  // served problem builds plans for many block widths)
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t bTasks = tdev.size() * sizeof(SolveTaskDev), bEnt = all.size() * sizeof(SolveEntry);
  const size_t bTQ = plan.slotTQ.size() * sizeof(int), bGrp = allGrp.size() * sizeof(GroupRec);
  const size_t oEnt = up(bTasks), oTQ = oEnt + up(std::max<size_t>(bEnt, 1)), oGrp = oTQ + up(bTQ);
  const size_t total = oGrp + up(std::max<size_t>(bGrp, 1));
  plan.oEnt = oEnt;
  plan.oTQ = oTQ;
  plan.oGrp = oGrp;
  if (plan.hostOnly) {
    std::vector<unsigned char> blob(total, 0);
    std::memcpy(blob.data(), tdev.data(), bTasks);
    if (bEnt) std::memcpy(blob.data() + oEnt, all.data(), bEnt);
    std::memcpy(blob.data() + oTQ, plan.slotTQ.data(), bTQ);
    if (bGrp) std::memcpy(blob.data() + oGrp, allGrp.data(), bGrp);
    plan.hBlob = std::move(blob);
    return;
  }
  // (the four parts uploaded from where they lie; the gaps are never read)
  LSW_HIP(devMalloc(&plan.dBlob, total));
  auto* dst = static_cast<unsigned char*>(plan.dBlob);
  LSW_HIP(uploadSync(dst, tdev.data(), bTasks));
  if (bEnt) LSW_HIP(uploadSync(dst + oEnt, all.data(), bEnt));
  LSW_HIP(uploadSync(dst + oTQ, plan.slotTQ.data(), bTQ));
  if (bGrp) LSW_HIP(uploadSync(dst + oGrp, allGrp.data(), bGrp));
  plan.bind(plan.dBlob);
  if (ptime)
    std::fprintf(stderr, "buildPlan tasks %d entries %zu groups %zu: entries_ms %.2f sort_ms %.2f launch_ms %.2f upload_ms %.2f\n",
                 nt, all.size(), allGrp.size(), 1e3 * (pt1 - pt0), 1e3 * (pt2 - pt1), 1e3 * (pt3 - pt2), 1e3 * (clk() - pt3));
}

void SolvePlan::bind(void* base) {
  auto* b = static_cast<unsigned char*>(base);
  dTasks = reinterpret_cast<SolveTaskDev*>(b);
  dEntries = reinterpret_cast<SolveEntry*>(b + oEnt);
  dSlotTQ = reinterpret_cast<int*>(b + oTQ);
  dGroups = reinterpret_cast<GroupRec*>(b + oGrp);
}

namespace {
struct LaunchTimer {
  SolveProfile* prof;
  hipStream_t s;
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
  std::vector<double> bytes, flops;
  double flopPeak = kPeakMfmaFp32;
  std::vector<const SolvePlan::Launch*> what;
  void begin() {
    if (!prof) return;
    hipEvent_t e;
    LSW_HIP(hipEventCreate(&e));
    LSW_HIP(hipEventRecord(e, s));
    ev.push_back(e);
  }
  void end(int k, double b, double f) {
    if (!prof) return;
    hipEvent_t e;
    LSW_HIP(hipEventCreate(&e));
    LSW_HIP(hipEventRecord(e, s));
    ev.push_back(e);
    kind.push_back(k);
    bytes.push_back(b);
    flops.push_back(f);
  }
  void finish() {
    if (!prof) return;
    LSW_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < kind.size(); ++i) {
      float ms = 0;
      LSW_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      if (std::getenv("LSW_PROFILE_LOG") && i < what.size()) {
        const auto& L = *what[i];
        std::fprintf(stderr, "launch %s %zu: entries %d mode %d rp %d wpb %d waves %d cols %d bytes %.0f us %.2f GB/s %.0f "
                     "flops %.0f ideal_us %.2f\n",
                     kind[i] ? "bwd" : "fwd", i, L.n, L.mode, L.rp, L.wpb, L.waves, L.cols, L.bytes, 1e3 * ms,
                     ms > 0 ? L.bytes / (ms * 1e6) : 0.0, L.flops,
                     1e6 * std::max(L.bytes / kPeakHbmBytesPerS, L.flops / flopPeak));
      }
      prof->idealMs += 1e3 * std::max(bytes[i] / kPeakHbmBytesPerS, flops[i] / flopPeak);
      if (kind[i] == 0) {
        prof->fwdMs += ms;
        prof->fwdLaunches++;
        prof->fwdBytes += bytes[i];
        prof->fwdFlops += flops[i];
      } else {
        prof->bwdMs += ms;
        prof->bwdLaunches++;
        prof->bwdBytes += bytes[i];
        prof->bwdFlops += flops[i];
      }
    }
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

struct SolveArgs {
  const SolveTaskDev* tasks;
  const SolveNode* nodes;
  const SolveEntry* ent;
  const GroupRec* grp;
  const int *poolE, *poolB, *poolMap;
  long long xs, gs;
};

template <typename T2, int NR, bool FWD>
void launch(const SolveArgs& a, const SolvePlan::Launch& L, const T2* factor, const T2* X, T2* Xo, T2* R, T2* G,
            hipStream_t s) {
  const SolveEntry* ent = a.ent + L.off;
  if (L.mode == 4) {
    if constexpr (std::is_same<T2, float2>::value && NR == kMaxRhs) {
      const unsigned grid = (unsigned)((L.ngroups + kGrpWaves - 1) / kGrpWaves);
      if (L.one)
        hipLaunchKernelGGL((k_nd_mfma_grp<FWD, true>), dim3(grid), dim3(64 * kGrpWaves), 0, s, a.grp + L.goff,
                           L.ngroups, a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
      else
        hipLaunchKernelGGL((k_nd_mfma_grp<FWD, false>), dim3(grid), dim3(64 * kGrpWaves), 0, s, a.grp + L.goff,
                           L.ngroups, a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
    } else {
      LSW_REQUIRE(false, LSW_ERR_STATE, "grouped matrix-core solve launch for a non-FP32 or narrow task plan");
    }
    return;
  }
  if (L.mode == 2) {
    if constexpr (std::is_same<T2, float2>::value && NR == kMaxRhs) {
      const int P = 64 / L.rp;
      const long long perBlock = (long long)P * L.wpb;
      const unsigned grid = (unsigned)((L.n + perBlock - 1) / perBlock);
      const size_t shm = (size_t)L.wpb * ((size_t)P * NR * L.cols + kMfmaOut) * sizeof(T2);
#define LSW_MFMA_LAUNCH(PR)                                                                                   \
  hipLaunchKernelGGL((k_nd_mfma<FWD, PR>), dim3(grid), dim3(64 * L.wpb), shm, s, a.tasks, a.nodes, ent, L.n,     \
                     ceilLog2(L.rp), L.cols, a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs)
      switch (L.pair) {
        case 1: LSW_MFMA_LAUNCH(1); break;
        case 2: LSW_MFMA_LAUNCH(2); break;
        default: LSW_MFMA_LAUNCH(0); break;
      }
#undef LSW_MFMA_LAUNCH
    } else {
      LSW_REQUIRE(false, LSW_ERR_STATE, "matrix-core solve launch for a non-FP32 or narrow task plan");
    }
    return;
  }
  if (L.mode == 3) {
    if constexpr (std::is_same<T2, float2>::value && NR == kMaxRhs) {
      const size_t shm = std::max((size_t)NR * L.cols * sizeof(T2), (size_t)L.waves * 64 * kMfmaRow * sizeof(float));
      const dim3 grid((unsigned)L.n);
      switch (L.waves) {
        case 1:
          hipLaunchKernelGGL((k_nd_mfma_wide<FWD, 1>), grid, dim3(64), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                             a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
          break;
        case 2:
          hipLaunchKernelGGL((k_nd_mfma_wide<FWD, 2>), grid, dim3(128), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                             a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
          break;
        case 4:
          hipLaunchKernelGGL((k_nd_mfma_wide<FWD, 4>), grid, dim3(256), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                             a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
          break;
        case 8:
          hipLaunchKernelGGL((k_nd_mfma_wide<FWD, 8>), grid, dim3(512), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                             a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
          break;
        default:
          LSW_REQUIRE(false, LSW_ERR_STATE, "matrix-core wide solve launch of an unsupported wavefront count");
      }
    } else {
      LSW_REQUIRE(false, LSW_ERR_STATE, "matrix-core wide solve launch for a non-FP32 or narrow task plan");
    }
    return;
  }
  if (L.mode == 0) {
    const int P = 64 / L.rp;
    const long long perBlock = (long long)P * L.rounds * L.wpb;
    const unsigned grid = (unsigned)((L.n + perBlock - 1) / perBlock);
    const size_t shm = (size_t)L.wpb * P * NR * L.cols * sizeof(T2);
    const int rpLog = ceilLog2(L.rp);
    hipLaunchKernelGGL((k_nd_packed<T2, NR, FWD>), dim3(grid), dim3(64 * L.wpb), shm, s, a.tasks, a.nodes, ent, L.n,
                       L.rounds, rpLog, L.cols, a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
    return;
  }
  const size_t shm = ((size_t)NR * L.cols + (size_t)L.waves * 64 * NR) * sizeof(T2);
  const dim3 grid((unsigned)L.n);
  switch (L.waves) {
    case 1:
      hipLaunchKernelGGL((k_nd_wide<T2, NR, 1, FWD>), grid, dim3(64), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                         a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
      break;
    case 2:
      hipLaunchKernelGGL((k_nd_wide<T2, NR, 2, FWD>), grid, dim3(128), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                         a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
      break;
    case 4:
      hipLaunchKernelGGL((k_nd_wide<T2, NR, 4, FWD>), grid, dim3(256), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                         a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
      break;
    case 8:
      hipLaunchKernelGGL((k_nd_wide<T2, NR, 8, FWD>), grid, dim3(512), shm, s, a.tasks, a.nodes, ent, L.n, L.cols,
                         a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
      break;
    default:
      // 16 wavefronts (1024 threads, at most 128 VGPRs): FP32 tasks of up to 4
      // right-hand sides only; the others would spill
      if constexpr (std::is_same<T2, float2>::value && NR <= kTaskRhs4)
        hipLaunchKernelGGL((k_nd_wide<T2, NR, 16, FWD>), grid, dim3(1024), shm, s, a.tasks, a.nodes, ent, L.n,
                           L.cols, a.poolE, a.poolB, a.poolMap, factor, X, Xo, R, G, a.xs, a.gs);
      else
        LSW_REQUIRE(false, LSW_ERR_STATE, "16-wavefront solve launch for FP64 or 8-right-hand-side tasks");
      break;
  }
}

template <typename T2, int NR>
void runPlan(const SolvePlan& plan, const SolveArgs& a, const T2* factor, const T2* X, T2* Xo, T2* R, T2* G,
             hipStream_t s, SolveProfile* prof) {
  LaunchTimer tm{prof, s, {}, {}, {}};
  tm.flopPeak = std::is_same<T2, double2>::value ? kPeakMfmaFp64 : kPeakMfmaFp32;
  for (const auto& L : plan.fwd) {
    tm.begin();
    launch<T2, NR, true>(a, L, factor, X, Xo, R, G, s);
    tm.end(0, L.bytes, L.flops);
    if (prof) tm.what.push_back(&L);
  }
  for (const auto& L : plan.bwd) {
    tm.begin();
    launch<T2, NR, false>(a, L, factor, X, Xo, R, G, s);
    tm.end(1, L.bytes, L.flops);
    if (prof) tm.what.push_back(&L);
  }
  LSW_HIP(hipGetLastError());
  tm.finish();
}
}  // namespace

void FactorSet::solve(const SolvePlan& plan, void* X, void* R, void* G, long long xs, long long gs,
                      hipStream_t s, SolveProfile* prof) const {
  solve(plan, X, X, R, G, xs, gs, s, prof);
}

void FactorSet::solve(const SolvePlan& plan, const void* X, void* Xo, void* R, void* G, long long xs, long long gs,
                      hipStream_t s, SolveProfile* prof) const {
  LSW_REQUIRE(mFactorized, LSW_ERR_STATE, "solve called before factorize");
  LSW_REQUIRE(Xo != X || !plan.outOfPlace, LSW_ERR_ARG, "plan built for separate right-hand side and solution buffers");
  if (plan.ntasks <= 0) return;
  const SolveArgs a{plan.dTasks, dSolveNodes, plan.dEntries, plan.dGroups, dPoolE, dPoolB, dPoolMap, xs, gs};
  if (mPrecision == LSW_FP64) {
    const auto* x = (const double2*)X;
    auto* xo = (double2*)Xo;
    switch (taskWidth(plan.maxNrhs)) {
      case 1: runPlan<double2, 1>(plan, a, dFactor, x, xo, (double2*)R, (double2*)G, s, prof); break;
      case 4: runPlan<double2, 4>(plan, a, dFactor, x, xo, (double2*)R, (double2*)G, s, prof); break;
      default: runPlan<double2, 8>(plan, a, dFactor, x, xo, (double2*)R, (double2*)G, s, prof); break;
    }
  } else {
    const auto* x = (const float2*)X;
    auto* xo = (float2*)Xo;
    switch (taskWidth(plan.maxNrhs)) {
      case 1: runPlan<float2, 1>(plan, a, dFactor32, x, xo, (float2*)R, (float2*)G, s, prof); break;
      case 4: runPlan<float2, 4>(plan, a, dFactor32, x, xo, (float2*)R, (float2*)G, s, prof); break;
      default: runPlan<float2, 8>(plan, a, dFactor32, x, xo, (float2*)R, (float2*)G, s, prof); break;
    }
  }
}

int FactorSet::maxTaskRhs() const {
  const double S = mPrecision == LSW_FP32 ? 8.0 : 16.0;
  // packed staging of one 64-row entry: 8 x front; wide: plus 2 wavefronts' partials
  return S * kMaxRhs * ((double)mMaxFront + 129.0) <= kMaxLds ? kMaxRhs : kTaskRhs4;
}

}  // namespace lsw
