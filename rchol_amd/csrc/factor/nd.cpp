// Nested-dissection orderings with the contract of the reference's
// find_separator (c++/rchol/find_separator.cpp:109-163): p lists the
// vertices as [left subtree, right subtree, separator] recursively, leaves
// keep increasing original index (:111-121), and the block sizes come in the
// same post-order (:138-142).  METIS is not used: grids are split
// geometrically, general graphs by breadth-first level structures.
#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../../include/rchol_factor.h"

namespace {

thread_local std::string g_nd_err;

struct Box {
  int lo[3], hi[3];  // half-open extents
  size_t count() const {
    size_t c = 1;
    for (int a = 0; a < 3; a++) c *= (size_t)std::max(0, hi[a] - lo[a]);
    return c;
  }
};

struct GridND {
  int nx, ny, nz, target;
  std::vector<size_t> p, val;
  bool ok = true;

  void emit(const Box &b) {
    for (int z = b.lo[2]; z < b.hi[2]; z++)
      for (int y = b.lo[1]; y < b.hi[1]; y++)
        for (int x = b.lo[0]; x < b.hi[0]; x++)
          p.push_back((size_t)x + (size_t)y * nx + (size_t)z * nx * ny);
  }

  void rec(const Box &b, int depth) {
    if (depth == target) {
      size_t before = p.size();
      emit(b);
      val.push_back(p.size() - before);
      return;
    }
    if (b.count() <= 1) {  // find_separator.cpp:122-126
      ok = false;
      return;
    }
    int ax = 0;
    for (int a = 1; a < 3; a++)
      if (b.hi[a] - b.lo[a] > b.hi[ax] - b.lo[ax]) ax = a;
    const int mid = b.lo[ax] + (b.hi[ax] - b.lo[ax]) / 2;
    Box L = b, R = b, S = b;
    L.hi[ax] = mid;
    R.lo[ax] = mid + 1;
    S.lo[ax] = mid;
    S.hi[ax] = mid + 1;
    rec(L, depth + 1);
    if (!ok) return;
    rec(R, depth + 1);
    if (!ok) return;
    size_t before = p.size();
    emit(S);
    val.push_back(p.size() - before);
  }
};

struct Graph {
  std::vector<size_t> rp, ci;  // local CSR without the diagonal
  std::vector<size_t> ids;     // local -> original index
};

// Vertex separator from a breadth-first level structure rooted at a
// pseudo-peripheral vertex: the level that halves the vertex count, then
// thinned (a separator vertex with no neighbour on the far side moves to the
// near side).  part[v]: 0 left, 1 right, 2 separator.
void bisect(const Graph &g, std::vector<int> &part) {
  const size_t n = g.ids.size();
  part.assign(n, 1);
  std::vector<long> lev(n, -1);
  std::vector<size_t> queue;
  auto bfs = [&](size_t s) {
    std::fill(lev.begin(), lev.end(), -1);
    queue.clear();
    queue.push_back(s);
    lev[s] = 0;
    for (size_t h = 0; h < queue.size(); h++) {
      size_t v = queue[h];
      for (size_t k = g.rp[v]; k < g.rp[v + 1]; k++) {
        size_t u = g.ci[k];
        if (lev[u] < 0) {
          lev[u] = lev[v] + 1;
          queue.push_back(u);
        }
      }
    }
    return queue.back();
  };
  size_t s = bfs(0);
  s = bfs(s);
  s = bfs(s);
  (void)s;
  const long maxl = lev[queue.back()];
  std::vector<size_t> cnt(maxl + 2, 0);
  for (size_t v = 0; v < n; v++)
    if (lev[v] >= 0) cnt[lev[v]]++;
  // first level whose prefix reaches half of the component
  const size_t comp = queue.size();
  size_t acc = 0;
  long L = 0;
  for (long l = 0; l <= maxl; l++) {
    if (acc + cnt[l] >= comp / 2) {
      L = l;
      break;
    }
    acc += cnt[l];
  }
  if (L == 0 && maxl >= 2) L = 1;
  for (size_t v = 0; v < n; v++) {
    if (lev[v] < 0) part[v] = 1;  // other components go right
    else if (lev[v] < L) part[v] = 0;
    else if (lev[v] == L) part[v] = 2;
    else part[v] = 1;
  }
  // thin the separator: a vertex without a right neighbour joins the left
  for (size_t v = 0; v < n; v++) {
    if (part[v] != 2) continue;
    bool touches_right = false;
    for (size_t k = g.rp[v]; k < g.rp[v + 1] && !touches_right; k++)
      touches_right = part[g.ci[k]] == 1;
    if (!touches_right) part[v] = 0;
  }
}

Graph induced(const Graph &g, const std::vector<int> &part, int which) {
  const size_t n = g.ids.size();
  std::vector<long> loc(n, -1);
  Graph s;
  for (size_t v = 0; v < n; v++)
    if (part[v] == which) {
      loc[v] = (long)s.ids.size();
      s.ids.push_back(g.ids[v]);
    }
  s.rp.push_back(0);
  for (size_t v = 0; v < n; v++) {
    if (part[v] != which) continue;
    for (size_t k = g.rp[v]; k < g.rp[v + 1]; k++)
      if (loc[g.ci[k]] >= 0) s.ci.push_back((size_t)loc[g.ci[k]]);
    s.rp.push_back(s.ci.size());
  }
  return s;
}

bool graph_rec(const Graph &g, int depth, int target, std::vector<size_t> &p,
               std::vector<size_t> &val) {
  if (depth == target) {
    p.insert(p.end(), g.ids.begin(), g.ids.end());
    val.push_back(g.ids.size());
    return true;
  }
  if (g.ids.size() <= 1) return false;
  std::vector<int> part;
  bisect(g, part);
  Graph l = induced(g, part, 0), r = induced(g, part, 1);
  if (!graph_rec(l, depth + 1, target, p, val)) return false;
  if (!graph_rec(r, depth + 1, target, p, val)) return false;
  size_t ns = 0;
  for (size_t v = 0; v < g.ids.size(); v++)
    if (part[v] == 2) {
      p.push_back(g.ids[v]);
      ns++;
    }
  val.push_back(ns);
  return true;
}

}  // namespace

extern "C" {

int rchol_nd_grid(int nx, int ny, int nz, int depth, size_t *perm, size_t *blocks) {
  if (nx < 1 || ny < 1 || nz < 1 || depth < 1 || depth > 30 || !perm || !blocks) return -1;
  GridND g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.target = depth;
  Box b{{0, 0, 0}, {nx, ny, nz}};
  g.rec(b, 1);
  if (!g.ok) return -3;  // too many leaves for this grid
  std::copy(g.p.begin(), g.p.end(), perm);
  std::copy(g.val.begin(), g.val.end(), blocks);
  return 0;
}

int rchol_nd_graph(size_t n, const size_t *rowPtr, const size_t *colIdx, int depth, size_t *perm,
                   size_t *blocks) {
  if (n == 0 || !rowPtr || !colIdx || depth < 1 || depth > 30 || !perm || !blocks) return -1;
  Graph g;
  g.ids.resize(n);
  g.rp.push_back(0);
  for (size_t v = 0; v < n; v++) {
    g.ids[v] = v;
    for (size_t k = rowPtr[v]; k < rowPtr[v + 1]; k++) {
      if (colIdx[k] >= n) return -1;
      if (colIdx[k] != v) g.ci.push_back(colIdx[k]);
    }
    g.rp.push_back(g.ci.size());
  }
  std::vector<size_t> p, val;
  if (!graph_rec(g, 1, depth, p, val)) return -3;
  std::copy(p.begin(), p.end(), perm);
  std::copy(val.begin(), val.end(), blocks);
  return 0;
}

}  // extern "C"
