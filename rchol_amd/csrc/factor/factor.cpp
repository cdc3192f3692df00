// Randomized Cholesky factorization of an SDDM matrix (host side of the
// solve engine).  Produces G (lower triangular, stored as the reference's
// SparseCSR G: CSC arrays) and the ordering that the GPU sweeps consume.
//
// The elimination restates the reference algorithm (rchol_lap.cpp):
//  * the SDDM matrix becomes a Laplacian with one extra "ground" vertex that
//    takes each row's excess (util.hpp:93-142), stored column-wise with the
//    diagonal first and edge weights positive (util.hpp:85-91, 60-83);
//  * columns are eliminated in order; fill appended to a column is sorted
//    by row and duplicates are summed (linear_update, rchol_lap.cpp:733-770);
//  * a column's neighbours are sorted by weight and each one but the last is
//    joined to one heavier neighbour drawn proportionally to the remaining
//    weight, with weight w_i (W - C_i) / W (random_sampling0, :643-729);
//  * the diagonal d = sum of neighbour weights; G(:,i) = [sqrt(d),
//    -w_j / sqrt(d)] without the ground row (coalesce, :358-387);
//  * the nested-dissection tree runs leaves in parallel; fill whose lower
//    endpoint leaves the current block travels to the parent, which merges
//    the left child's list then the right child's before its separator
//    (recursive_calculation, :187-355).
// Random numbers: one std::mt19937(seed) per "spine" of the tree (the
// thread that runs a node's left child runs the node's separator too; a right
// child starts a new spine), which is exactly the per-thread generator layout
// of the reference when its std::random_device yields `seed`.
//
// ORDER_MIS additionally chooses the order *inside* each block while it is
// eliminated: repeatedly take a maximal independent set of the block's
// current elimination graph (lowest dependency level first) and eliminate
// it.  Vertices of one set share no edge, so each set is one level of the
// triangular sweeps; the sweep depth drops from ~block size to ~number of
// sets.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/rchol_factor.h"
#include "parallel.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct Ent {
  size_t row;
  double w;
};
struct ByRow {
  bool operator()(const Ent &a, const Ent &b) const { return a.row < b.row; }
};
struct ByWeight {
  bool operator()(const Ent &a, const Ent &b) const { return a.w < b.w; }
};
struct SepEdge {
  double w;
  size_t row;  // larger endpoint
  size_t col;  // smaller endpoint
};

struct Column {
  std::vector<Ent> e;
  size_t split = 0;  // entries present before any fill
};

struct Spine {
  std::mt19937 gen;
  explicit Spine(unsigned int seed) : gen(seed) {}
};

class Factorizer {
 public:
  size_t n = 0;     // real vertices; the ground vertex is n
  std::vector<Column> cols;
  std::vector<double> diag;
  std::vector<size_t> ri;  // block boundaries (result_idx)
  unsigned int seed = 0;
  int target = 1;
  int order = RCHOL_ORDER_STATIC;
  // MIS mode: final position of every vertex (ND numbering -> position)
  std::vector<size_t> pos;
  std::unique_ptr<std::atomic<int>[]> level;  // dependency level estimate (MIS mode)

  int lev(size_t v) const { return level[v].load(std::memory_order_relaxed); }
  void raise_level(size_t v, int l) {
    int cur = level[v].load(std::memory_order_relaxed);
    while (cur < l && !level[v].compare_exchange_weak(cur, l, std::memory_order_relaxed)) {
    }
  }

  // --- static elimination of one column (rchol_lap.cpp:733-770, 643-729) ---
  static void merge_sorted(Column &col) {
    std::vector<Ent> s(col.e);
    std::sort(s.begin(), s.end(), ByRow());
    size_t add = 0;
    size_t rowp = s[0].row;
    col.e[0] = s[0];
    for (size_t i = 1; i < s.size(); i++) {
      if (s[i].row != rowp) {
        add++;
        rowp = s[i].row;
        col.e[add] = s[i];
      } else {
        col.e[add].w += s[i].w;
      }
    }
    col.e.resize(add + 1);
  }

  // Samples the clique of column c.  Entries e[1..] are its neighbours (any
  // order that is deterministic); emits fill via emit(row_lo, row_hi, w).
  template <typename Emit>
  static double sample_clique(std::vector<Ent> &e, Spine &sp, Emit emit) {
    const size_t nz = e.size();
    double sum = 0.0;
    for (size_t i = 1; i < nz; i++) sum += e[i].w;
    if (nz > 2) {
      const size_t size = nz - 1;
      std::vector<Ent> s(e.begin() + 1, e.end());
      std::sort(s.begin(), s.end(), ByWeight());
      std::vector<double> cumsum(size);
      double csum = 0.0;
      for (size_t i = 0; i < size; i++) {
        csum += s[i].w;
        cumsum[i] = csum;
      }
      for (size_t i = 0; i + 1 < size; i++) {
        std::uniform_real_distribution<> dis(0.0, 1.0);
        const double tar = cumsum[i];
        const double r = dis(sp.gen) * (csum - tar) + tar;
        size_t low = std::lower_bound(cumsum.begin() + i, cumsum.end(), r) - cumsum.begin();
        if (low >= size) low = size - 1;  // r rounded above the total
        const size_t a = s[low].row, b = s[i].row;
        emit(std::min(a, b), std::max(a, b), s[i].w * (csum - cumsum[i]) / csum);
      }
    }
    const double inv = 1.0 / sum;
    for (size_t i = 0; i < nz; i++) e[i].w *= inv;
    e[0].w = 1.0;
    return sum;
  }

  void eliminate_static(size_t c, Spine &sp, size_t lb, size_t rb, std::vector<SepEdge> &sep) {
    Column &col = cols[c];
    if (col.e.size() > col.split) merge_sorted(col);
    diag[c] = sample_clique(col.e, sp, [&](size_t lo, size_t hi, double w) {
      if (lo >= lb && lo < rb)
        cols[lo].e.push_back({hi, w});
      else
        sep.push_back({w, hi, lo});
    });
  }

  // --- MIS ordering inside a block -----------------------------------------
  // Vertices of the block are [lb, rb) in ND numbering; the block's
  // adjacency is made symmetric for the vertices inside it, eliminated in
  // independent-set layers, and the block's positions [lb, rb) are handed
  // out in elimination order.
  void eliminate_block_mis(size_t lb, size_t rb, Spine &sp, std::vector<SepEdge> &sep) {
    const size_t m = std::min(rb, n) - lb;  // the ground vertex is never eliminated here
    if (m == 0) return;
    // adj[v - lb]: current neighbours of block vertex v (may hold eliminated
    // vertices and duplicates; both are dropped when v is eliminated).
    std::vector<std::vector<Ent>> adj(m);
    for (size_t v = lb; v < lb + m; v++) {
      Column &col = cols[v];
      for (size_t k = 0; k < col.e.size(); k++) {
        const Ent &x = col.e[k];
        if (x.row == v) continue;  // diagonal
        adj[v - lb].push_back(x);
        if (x.row >= lb && x.row < lb + m) adj[x.row - lb].push_back({v, x.w});
      }
      col.e.clear();
      col.e.shrink_to_fit();
    }
    std::vector<char> alive(m, 1), blocked(m, 0);
    std::vector<size_t> cand(m);
    size_t remaining = m, next_pos = lb;
    std::vector<Ent> nb;
    while (remaining) {
      // candidates: alive vertices, lowest level first, then fewest neighbours
      size_t nc = 0;
      for (size_t i = 0; i < m; i++)
        if (alive[i]) cand[nc++] = i;
      std::sort(cand.begin(), cand.begin() + nc, [&](size_t a, size_t b) {
        const int la = lev(lb + a), lbv = lev(lb + b);
        if (la != lbv) return la < lbv;
        const size_t da = adj[a].size(), db = adj[b].size();
        if (da != db) return da < db;
        return a < b;
      });
      std::vector<size_t> layer;
      for (size_t t = 0; t < nc; t++) {
        const size_t i = cand[t];
        if (blocked[i]) continue;
        layer.push_back(i);
        for (const Ent &x : adj[i])
          if (x.row >= lb && x.row < lb + m && alive[x.row - lb]) blocked[x.row - lb] = 1;
      }
      for (size_t t = 0; t < nc; t++) blocked[cand[t]] = 0;
      // the layer is chosen in priority order but numbered (and eliminated:
      // its vertices are independent, so any order is valid) in ND order,
      // so that consecutive positions stay neighbours in the domain
      std::sort(layer.begin(), layer.end());
      for (size_t i : layer) {
        const size_t v = lb + i;
        // current neighbour list: live block vertices and outside vertices,
        // merged by vertex (ND numbering ascending)
        nb.clear();
        nb.push_back({v, 0.0});
        std::vector<Ent> &a = adj[i];
        std::sort(a.begin(), a.end(), ByRow());
        for (size_t k = 0; k < a.size(); k++) {
          const size_t u = a[k].row;
          if (u >= lb && u < lb + m && !alive[u - lb]) continue;
          if (nb.size() > 1 && nb.back().row == u)
            nb.back().w += a[k].w;
          else
            nb.push_back(a[k]);
        }
        std::vector<Ent>().swap(a);
        const int lv = lev(v) + 1;
        for (size_t k = 1; k < nb.size(); k++)
          if (nb[k].row < n) raise_level(nb[k].row, lv);
        diag[v] = sample_clique(nb, sp, [&](size_t lo, size_t hi, double w) {
          const bool lin = lo >= lb && lo < lb + m, hin = hi >= lb && hi < lb + m;
          if (lin) adj[lo - lb].push_back({hi, w});
          if (hin) adj[hi - lb].push_back({lo, w});
          if (!lin && !hin) sep.push_back({w, hi, lo});
        });
        cols[v].e = nb;  // column of G in ND numbering (diagonal first)
        alive[i] = 0;
        pos[v] = next_pos++;
        remaining--;
      }
    }
  }

  // --- the nested-dissection tree (rchol_lap.cpp:187-355) --------------------
  void run_range(size_t lb, size_t rb, Spine &sp, std::vector<SepEdge> &sep) {
    if (order == RCHOL_ORDER_MIS) {
      eliminate_block_mis(lb, rb, sp, sep);
      return;
    }
    for (size_t c = lb; c < rb; c++) {
      if (c == n) continue;  // ground column: no neighbours later, never emitted
      eliminate_static(c, sp, lb, rb, sep);
    }
  }

  std::vector<SepEdge> recurse(size_t start, size_t total, int depth, Spine &sp) {
    std::vector<SepEdge> out;
    if (depth == target) {
      run_range(ri[start], ri[start + total], sp, out);
      return out;
    }
    const size_t half = (total - 1) / 2;
    std::vector<SepEdge> redges;
    std::exception_ptr rerr;
    std::thread t([&]() {
      try {
        Spine child(seed);
        redges = recurse(start + half, half, depth + 1, child);
      } catch (...) {
        rerr = std::current_exception();
      }
    });
    std::vector<SepEdge> ledges;
    try {
      ledges = recurse(start, half, depth + 1, sp);
    } catch (...) {
      t.join();
      throw;
    }
    t.join();
    if (rerr) std::rethrow_exception(rerr);
    const size_t lo = ri[start + total - 1], hi = ri[start + total];
    for (const std::vector<SepEdge> *lst : {&ledges, &redges}) {
      for (const SepEdge &x : *lst) {
        if (x.col >= lo && x.col < hi)
          cols[x.col].e.push_back({x.row, x.w});
        else
          out.push_back(x);
      }
    }
    std::vector<SepEdge>().swap(ledges);
    std::vector<SepEdge>().swap(redges);
    run_range(lo, hi, sp, out);
    return out;
  }
};

}  // namespace

extern "C" {

const char *rchol_factor_last_error(void) { return g_err.c_str(); }

void rchol_factor_free(void *p) { free(p); }

int rchol_factor(size_t n, const size_t *rowPtr, const size_t *colIdx, const double *val,
                 const size_t *perm, const size_t *blocks, size_t nblocks, int order,
                 unsigned int seed, rchol_factor_out *out) {
  if (!rowPtr || !colIdx || !val || !blocks || !out) return fail(-1, "null argument");
  if (n == 0) return fail(-1, "empty matrix");
  if (order != RCHOL_ORDER_STATIC && order != RCHOL_ORDER_MIS) return fail(-1, "unknown order");
  // nblocks = 2^d - 1 (rchol_parallel.cpp:62-70 / find_separator.cpp:138-150)
  size_t d = 0;
  while (((size_t)1 << (d + 1)) - 1 < nblocks) d++;
  if (((size_t)1 << (d + 1)) - 1 != nblocks) return fail(-1, "nblocks must be 2^d - 1");
  size_t total = 0;
  for (size_t i = 0; i < nblocks; i++) total += blocks[i];
  if (total != n) return fail(-1, "block sizes do not add up to n");
  try {
    Factorizer F;
    F.n = n;
    F.seed = seed;
    F.order = order;
    F.target = (int)d + 1;
    F.ri.assign(1, 0);
    for (size_t i = 0; i < nblocks; i++) F.ri.push_back(F.ri.back() + blocks[i]);
    F.ri.back()++;  // the ground vertex joins the last block
    std::vector<size_t> p(n), transp(n);
    for (size_t i = 0; i < n; i++) p[i] = perm ? perm[i] : i;
    {
      std::vector<char> seen(n, 0);
      for (size_t i = 0; i < n; i++) {
        if (p[i] >= n || seen[p[i]]) return fail(-1, "perm is not a permutation");
        seen[p[i]] = 1;
        transp[p[i]] = i;
      }
    }
    // Columns of the Laplacian's lower triangle in the new numbering
    // (util.cpp:16-56 reorder + util.hpp:60-142 + rchol_lap.cpp:778-855).
    F.cols.resize(n + 1);
    F.diag.assign(n + 1, 0.0);
    {
      std::atomic<bool> bad{false};
      rchol_host::parallel_for(rowPtr[n], [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; k++)
          if (colIdx[k] >= n) {
            bad.store(true, std::memory_order_relaxed);
            return;
          }
      });
      if (bad.load()) return fail(-1, "column index out of range");
    }
    rchol_host::parallel_for(n, [&](size_t i0, size_t i1) {
      std::vector<Ent> row;
      for (size_t i = i0; i < i1; i++) {
        const size_t o = p[i];
        row.clear();
        for (size_t k = rowPtr[o]; k < rowPtr[o + 1]; k++) row.push_back({transp[colIdx[k]], val[k]});
        if (perm) std::sort(row.begin(), row.end(), ByRow());
        double sum = 0.0;
        for (const Ent &x : row) sum += x.w;
        Column &col = F.cols[i];
        size_t keep = 1;  // the ground entry, if any
        for (const Ent &x : row) keep += x.row >= i;
        col.e.reserve(keep);
        for (const Ent &x : row)
          if (x.row >= i) col.e.push_back({x.row, x.w * -1});
        if (std::abs(sum) > 1e-09) {
          const double g = (-sum) * -1;
          if (!(g < 0)) col.e.push_back({n, g});
        }
        col.split = col.e.size();
      }
    });
    F.pos.assign(n + 1, 0);
    F.level.reset(new std::atomic<int>[n + 1]);
    for (size_t i = 0; i <= n; i++) F.level[i].store(0, std::memory_order_relaxed);
    for (size_t i = 0; i <= n; i++) F.pos[i] = i;
    auto t0 = std::chrono::steady_clock::now();
    Spine root(seed);
    F.recurse(0, nblocks, 1, root);
    auto t1 = std::chrono::steady_clock::now();
    // coalesce (rchol_lap.cpp:358-387), rows mapped to final positions
    std::vector<size_t> col_at(n);
    for (size_t v = 0; v < n; v++) col_at[F.pos[v]] = v;
    std::vector<size_t> cnt(n + 1, 0);
    rchol_host::parallel_for(n, [&](size_t c0, size_t c1) {
      for (size_t c = c0; c < c1; c++) {
        size_t k = 0;
        for (const Ent &x : F.cols[col_at[c]].e) k += (x.row != n);
        cnt[c + 1] = k;
      }
    });
    for (size_t c = 0; c < n; c++) cnt[c + 1] += cnt[c];
    const size_t nnz = cnt[n];
    out->n = n;
    out->colPtr = (size_t *)malloc((n + 1) * sizeof(size_t));
    out->rowIdx = (size_t *)malloc(std::max<size_t>(nnz, 1) * sizeof(size_t));
    out->val = (double *)malloc(std::max<size_t>(nnz, 1) * sizeof(double));
    out->perm = (size_t *)malloc(n * sizeof(size_t));
    if (!out->colPtr || !out->rowIdx || !out->val || !out->perm) return fail(-2, "out of memory");
    std::copy(cnt.begin(), cnt.end(), out->colPtr);
    rchol_host::parallel_for(n, [&](size_t c0, size_t c1) {
      std::vector<Ent> tmp;
      for (size_t c = c0; c < c1; c++) {
        const size_t v = col_at[c];
        const std::vector<Ent> &e = F.cols[v].e;
        const double sd = std::sqrt(F.diag[v]);
        size_t at = cnt[c];
        if (order == RCHOL_ORDER_STATIC) {
          for (const Ent &x : e) {
            if (x.row == n) continue;
            double g = x.w;
            if (x.row != v) g = g * -1;
            out->rowIdx[at] = x.row;
            out->val[at] = g * sd;
            at++;
          }
        } else {
          tmp.clear();
          for (const Ent &x : e) {
            if (x.row == n) continue;
            double g = x.w;
            if (x.row != v) g = g * -1;
            tmp.push_back({F.pos[x.row], g * sd});
          }
          std::sort(tmp.begin(), tmp.end(), ByRow());
          for (const Ent &x : tmp) {
            out->rowIdx[at] = x.row;
            out->val[at] = x.w;
            at++;
          }
        }
      }
    });
    for (size_t v = 0; v < n; v++) out->perm[F.pos[v]] = p[v];
    out->seconds = std::chrono::duration<double>(t1 - t0).count();
    // release the n column lists on all threads (one thread freeing them in
    // F's destructor takes longer than the rest of the hand-off)
    rchol_host::parallel_for(F.cols.size(), [&](size_t c0, size_t c1) {
      for (size_t c = c0; c < c1; c++) std::vector<Ent>().swap(F.cols[c].e);
    });
  } catch (const std::exception &ex) {
    return fail(-2, std::string("factorization failed: ") + ex.what());
  }
  return 0;
}

}  // extern "C"
