// Nested-dissection symbolic analysis (see lsw_nd.h).
#include "lsw_nd.h"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

namespace lsw {

int leafDofs() {
  static const int v = [] {
    const char* e = std::getenv("LSW_LEAF_DOFS");
    const int x = e ? std::atoi(e) : 64;
    return std::max(4, std::min(256, x));
  }();
  return v;
}

namespace {
struct Region {
  int i0, i1, j0, j1;
};
}  // namespace

int NdTree::rec(int i0, int i1, int j0, int j1, int parent) {
  const int w = i1 - i0, h = j1 - j0;
  if (w <= 0 || h <= 0) return -1;
  const int id = (int)nodes.size();
  nodes.emplace_back();
  nodes[id].parent = parent;
  std::vector<int> Eloc;
  int c1 = -1, c2 = -1;
  if (w * h <= leafDofs()) {
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j) Eloc.push_back(i * ny + j);
  } else if (w >= h) {
    const int s = i0 + w / 2;
    for (int j = j0; j < j1; ++j) Eloc.push_back(s * ny + j);
    c1 = rec(i0, s, j0, j1, id);
    c2 = rec(s + 1, i1, j0, j1, id);
  } else {
    const int s = j0 + h / 2;
    for (int i = i0; i < i1; ++i) Eloc.push_back(i * ny + s);
    c1 = rec(i0, i1, j0, s, id);
    c2 = rec(i0, i1, s + 1, j1, id);
  }
  std::vector<int> Bloc;
  if (i0 > 0)
    for (int j = j0; j < j1; ++j) Bloc.push_back((i0 - 1) * ny + j);
  if (i1 < nx)
    for (int j = j0; j < j1; ++j) Bloc.push_back(i1 * ny + j);
  if (j0 > 0)
    for (int i = i0; i < i1; ++i) Bloc.push_back(i * ny + (j0 - 1));
  if (j1 < ny)
    for (int i = i0; i < i1; ++i) Bloc.push_back(i * ny + j1);
  std::sort(Bloc.begin(), Bloc.end());
  NdNode& nd = nodes[id];
  nd.i0 = i0;
  nd.i1 = i1;
  nd.j0 = j0;
  nd.j1 = j1;
  nd.c1 = c1;
  nd.c2 = c2;
  nd.e = (int)Eloc.size();
  nd.b = (int)Bloc.size();
  nd.eOff = (int)E.size();
  E.insert(E.end(), Eloc.begin(), Eloc.end());
  nd.bOff = (int)B.size();
  B.insert(B.end(), Bloc.begin(), Bloc.end());
  int hgt = 0;
  if (c1 >= 0) hgt = std::max(hgt, nodes[c1].height + 1);
  if (c2 >= 0) hgt = std::max(hgt, nodes[c2].height + 1);
  nodes[id].height = hgt;
  return id;
}

void NdTree::build(int nx_, int ny_) {
  nx = nx_;
  ny = ny_;
  n = nx * ny;
  nodes.clear();
  E.clear();
  B.clear();
  cmap.clear();
  up.clear();
  asmList.clear();
  rec(0, nx, 0, ny, -1);
  if ((int)E.size() != n) throw std::logic_error("ND: separators do not cover the grid");

  maxHeight = 0;
  for (auto& nd : nodes) maxHeight = std::max(maxHeight, nd.height);
  byHeight.assign(maxHeight + 1, {});
  for (int p = 0; p < (int)nodes.size(); ++p) byHeight[nodes[p].height].push_back(p);

  // front positions, child maps, up positions, assembly lists
  std::vector<int> pos(n, -1);
  long long fo = 0, go = 0, so = 0;
  maxFront = 0;
  for (int p = 0; p < (int)nodes.size(); ++p) {
    NdNode& nd = nodes[p];
    const int F = nd.e + nd.b;
    maxFront = std::max(maxFront, F);
    for (int k = 0; k < nd.e; ++k) pos[E[nd.eOff + k]] = k;
    for (int t = 0; t < nd.b; ++t) pos[B[nd.bOff + t]] = nd.e + t;
    // child maps
    nd.mapOff = (int)cmap.size();
    cmap.resize(cmap.size() + 2 * (size_t)F, -1);
    const int cs[2] = {nd.c1, nd.c2};
    for (int ci = 0; ci < 2; ++ci) {
      const int c = cs[ci];
      if (c < 0) continue;
      const NdNode& cn = nodes[c];
      for (int t = 0; t < cn.b; ++t) {
        const int d = B[cn.bOff + t];
        const int f = pos[d];
        if (f < 0) throw std::logic_error("ND: child boundary outside parent front");
        cmap[nd.mapOff + ci * F + f] = t;
      }
    }
    // assembly: rows of E (all stencil entries inside the front), then the
    // entries A[q][d] of boundary rows q in B coupling to d in E.
    nd.asmOff = (int)asmList.size();
    auto nb = [&](int d, int dir) -> int {
      const int i = d / ny, j = d % ny;
      switch (dir) {
        case 0: return d;
        case 1: return i > 0 ? d - ny : -1;
        case 2: return i < nx - 1 ? d + ny : -1;
        case 3: return j > 0 ? d - 1 : -1;
        default: return j < ny - 1 ? d + 1 : -1;
      }
    };
    static const int opposite[5] = {0, 2, 1, 4, 3};
    for (int k = 0; k < nd.e; ++k) {
      const int d = E[nd.eOff + k];
      for (int dir = 0; dir < 5; ++dir) {
        const int q = nb(d, dir);
        if (q < 0) continue;
        const int f = pos[q];
        if (f >= 0) asmList.push_back({k, f, d, dir});
      }
    }
    for (int t = 0; t < nd.b; ++t) {
      const int q = B[nd.bOff + t];
      for (int dir = 1; dir < 5; ++dir) {
        const int d = nb(q, dir);
        if (d < 0) continue;
        const int f = pos[d];
        if (f >= 0 && f < nd.e) asmList.push_back({nd.e + t, f, q, dir});
      }
      (void)opposite;
    }
    nd.asmCnt = (int)asmList.size() - nd.asmOff;
    // offsets
    nd.fwdOff = fo;
    fo += (long long)nd.b * nd.e;
    nd.bwdOff = fo;
    fo += (long long)nd.e * F;
    nd.gOff = (int)go;
    go += nd.b;
    nd.sOff = so;
    so += (long long)nd.b * nd.b;
    for (int k = 0; k < nd.e; ++k) pos[E[nd.eOff + k]] = -1;
    for (int t = 0; t < nd.b; ++t) pos[B[nd.bOff + t]] = -1;
  }
  // up positions (child B -> parent front position)
  for (int p = 0; p < (int)nodes.size(); ++p) {
    NdNode& nd = nodes[p];
    nd.upOff = (int)up.size();
    up.resize(up.size() + nd.b, -1);
  }
  for (int p = 0; p < (int)nodes.size(); ++p) {
    const NdNode& nd = nodes[p];
    const int F = nd.e + nd.b;
    const int cs[2] = {nd.c1, nd.c2};
    for (int ci = 0; ci < 2; ++ci) {
      const int c = cs[ci];
      if (c < 0) continue;
      for (int f = 0; f < F; ++f) {
        const int t = cmap[nd.mapOff + ci * F + f];
        if (t >= 0) up[nodes[c].upOff + t] = f;
      }
    }
  }
  for (int p = 0; p < (int)nodes.size(); ++p)
    for (int t = 0; t < nodes[p].b; ++t)
      if (up[nodes[p].upOff + t] < 0) throw std::logic_error("ND: boundary entry not mapped");
  factorSize = fo;
  gSize = go;
  schurSize = so;

  fwdItems.assign(maxHeight + 1, {});
  bwdItems.assign(maxHeight + 1, {});
  for (int h = 0; h <= maxHeight; ++h) {
    for (int p : byHeight[h]) {
      const NdNode& nd = nodes[p];
      if (nd.b == 0) {
        fwdItems[h].push_back({p, 0, 0, 0});
      } else {
        for (int r0 = 0; r0 < nd.b; r0 += kSolveRows)
          fwdItems[h].push_back({p, r0, std::min(kSolveRows, nd.b - r0), 0});
      }
      for (int r0 = 0; r0 < nd.e; r0 += kSolveRows)
        bwdItems[h].push_back({p, r0, std::min(kSolveRows, nd.e - r0), 0});
    }
  }
}

std::shared_ptr<const NdTree> sharedTree(int nx, int ny) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::weak_ptr<const NdTree>> cache;
  const auto key = std::make_pair(nx, ny);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end())
      if (auto t = it->second.lock()) return t;
  }
  // built without the lock (threads set up different shapes at once); when
  // two threads build the same shape, the first one stored is kept
  auto t = std::make_shared<NdTree>();
  t->build(nx, ny);
  std::lock_guard<std::mutex> lk(mu);
  auto& slot = cache[key];
  if (auto had = slot.lock()) return had;
  slot = t;
  if (cache.size() > 4096)
    for (auto it = cache.begin(); it != cache.end();) it = it->second.expired() ? cache.erase(it) : std::next(it);
  return t;
}

std::vector<NdNodeDev> NdTree::deviceNodes() const {
  std::vector<NdNodeDev> out(nodes.size());
  for (size_t p = 0; p < nodes.size(); ++p) {
    const NdNode& s = nodes[p];
    NdNodeDev& d = out[p];
    d.fwdOff = s.fwdOff;
    d.bwdOff = s.bwdOff;
    d.sOff = s.sOff;
    d.wsOff = s.wsOff;
    d.e = s.e;
    d.b = s.b;
    d.eOff = s.eOff;
    d.bOff = s.bOff;
    d.c1 = s.c1;
    d.c2 = s.c2;
    d.mapOff = s.mapOff;
    d.upOff = s.upOff;
    d.gOff = s.gOff;
    d.asmOff = s.asmOff;
    d.asmCnt = s.asmCnt;
    d.pad = 0;
  }
  return out;
}

}  // namespace lsw
