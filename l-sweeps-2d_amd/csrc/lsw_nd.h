// Nested-dissection symbolic analysis of the 5-point operator on an nx x ny
// grid (DOF index i*ny + j, y fastest -- the numbering of
// AssemblerHelmholtz2DFD::getDOFPts, src/AssemblerHelmholtz2DFD.cpp:147-181).
//
// The tree is geometric: a region is split by one grid line across its longer
// side until it holds at most leafDofs() DOFs.  Node p eliminates E_p (its
// separator line, or the whole region for a leaf); its boundary B_p is the
// set of grid points adjacent to the region, which all lie on ancestor
// separators.  The front of p is F_p = [E_p, B_p].
//
// Factor storage per node (complex, column-major):
//   W_p = F21 F11^{-1}                    b x e   (forward solve)
//   K_p = [F11^{-1} | -F11^{-1} F12]      e x (e+b) (backward solve)
// so both triangular solves become dense matrix-vector products.
#pragma once

#include <memory>
#include <vector>

namespace lsw {

// Largest leaf (DOFs).  Default 64 (an 8 x 8 block: one wavefront of rows);
// LSW_LEAF_DOFS overrides it (4..256) for tuning.  Larger leaves mean fewer
// tree heights, so fewer solve launches per sweep step, and denser leaf
// factors: on a 1600^2 grid of 8 x 8 cells, leaves of 16/32/64/128 DOFs gave
// 57.7/53.2/52.0/53.7 ms per FP32-factor solve.
int leafDofs();
constexpr int kSolveRows = 64;  // rows per solve work item (one wavefront)

struct NdNode {
  int i0 = 0, i1 = 0, j0 = 0, j1 = 0;  // region [i0,i1) x [j0,j1) of the subtree
  int e = 0, b = 0;
  int eOff = 0, bOff = 0;   // offsets into E / B pools
  int c1 = -1, c2 = -1, parent = -1;
  int height = 0;
  int mapOff = 0;           // m1 then m2, each e+b ints: front pos -> child g index or -1
  int upOff = 0;            // b ints: position of B_p[t] in the parent's front
  int gOff = 0;             // offset of g_p (b entries) in a task's update buffer
  int asmOff = 0, asmCnt = 0;
  long long fwdOff = 0, bwdOff = 0;  // W_p and K_p offsets in the cell's factor blob
  long long sOff = 0;                // Schur complement (b*b) offset in a cell's workspace
  long long wsOff = 0;               // dense front workspace offset (within a height batch)
};

// Device-side mirror (POD, 16-byte aligned fields first).
struct NdNodeDev {
  long long fwdOff, bwdOff, sOff, wsOff;
  int e, b, eOff, bOff, c1, c2, mapOff, upOff, gOff, asmOff, asmCnt, pad;
};

// One original matrix entry scattered into a front: F[fr][fc] += A[dof][dir].
struct AsmEntry {
  int fr, fc, dof, dir;
};

// Solve work item: rows [r0, r0+rows) of node `node` (forward: rows of W,
// backward: rows of K).
struct SolveItem {
  int node, r0, rows, pad;
};

struct NdTree {
  int nx = 0, ny = 0, n = 0;
  std::vector<NdNode> nodes;
  std::vector<int> E, B, cmap, up;
  std::vector<AsmEntry> asmList;
  int maxHeight = 0;
  long long factorSize = 0;  // complex entries per cell
  long long gSize = 0;       // update-vector entries per right-hand side
  long long schurSize = 0;   // Schur workspace entries per cell
  int maxFront = 0;          // max e+b
  std::vector<std::vector<int>> byHeight;
  std::vector<std::vector<SolveItem>> fwdItems, bwdItems;  // per height

  void build(int nx_, int ny_);
  std::vector<NdNodeDev> deviceNodes() const;

 private:
  int rec(int i0, int i1, int j0, int j1, int parent);
};

// The tree of an nx x ny grid.  A tree never changes once built, so grids of
// one shape share it: a process-wide cache of weak references hands out the
// tree while any holder keeps it (a problem whose memory estimate built its
// trees at registration reuses them when it is set up on the device).
std::shared_ptr<const NdTree> sharedTree(int nx, int ny);

}  // namespace lsw
