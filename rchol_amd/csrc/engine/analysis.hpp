// Host analysis of a stored factor for the triangular sweeps: dependency
// levels of the forward and the backward sweep, rows in level ("slot")
// order and their off-diagonal entries in that order.  Plain C++ (no HIP):
// included by engine.hip inside its anonymous namespace, after fail() and
// kTaskRows, and after <algorithm>, <atomic>, <chrono>, <memory>, <mutex>,
// <thread>, <vector> and <sys/mman.h>.
//
// Every pass is parallel over host threads and walks the factor in its own
// (nested-dissection) order, so the level reads and the scattered writes of
// a thread stay inside the piece of the domain it works on:
//   1. columns: validation, pivots, entry counts per row bucket, and which
//      pieces of rows depend on which;
//   2. the CSR of lower G in natural row order (a transposition through
//      row buckets, with no atomic operations);
//   3. levels of both sweeps over pieces of rows, each piece a task that
//      runs once every piece it reads is done; each row's entries sorted
//      for its sweep (backward rows inside those tasks, forward rows after);
//   4. slots (stable counting sort by level).
// Rows stay in natural order, each sorted for its sweep in step 3; the
// device gathers them into slot order.
// The result does not depend on the thread count.
#pragma once

// Allocator whose resize leaves new elements uninitialised: the analysis
// writes every element of the entry arrays from several threads, and a
// zero-fill first would be a serial pass over gigabytes.  Allocations of
// 64 MiB and more are 2 MiB aligned and advised for transparent huge
// pages, which makes the threads' first touch of them cheaper.
constexpr size_t kHugeAllocBytes = 64ull << 20;
constexpr size_t kHugePage = 2ull << 20;

template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <typename U>
  NoInit(const NoInit<U> &) {}
  T *allocate(size_t k) {
    const size_t bytes = k * sizeof(T);
    if (bytes < kHugeAllocBytes) return std::allocator<T>::allocate(k);
    void *p = nullptr;
    if (posix_memalign(&p, kHugePage, bytes)) throw std::bad_alloc();
    (void)madvise(p, (bytes + kHugePage - 1) & ~(kHugePage - 1), MADV_HUGEPAGE);
    return (T *)p;
  }
  void deallocate(T *p, size_t k) {
    if (k * sizeof(T) < kHugeAllocBytes)
      std::allocator<T>::deallocate(p, k);
    else
      free(p);
  }
  template <typename U>
  void construct(U *p) noexcept {
    ::new ((void *)p) U;
  }
  template <typename U, typename... Args>
  void construct(U *p, Args &&...args) {
    ::new ((void *)p) U(std::forward<Args>(args)...);
  }
};
template <typename T, typename U>
bool operator==(const NoInit<T> &, const NoInit<U> &) {
  return true;
}
template <typename T, typename U>
bool operator!=(const NoInit<T> &, const NoInit<U> &) {
  return false;
}

using IntVec = std::vector<int, NoInit<int>>;

using DblVec = std::vector<double, NoInit<double>>;

// Host-side schedule of one sweep before upload.  The rows' entries are
// kept in natural row order (row r at [nat_rp[r], nat_rp[r + 1]) of nat_ci /
// nat_cv, each row already in sweep order, in `store`); the device gathers
// them into slot order, and the chunk planner reads them as they are.
struct HostSweep {
  std::vector<int> task_ptr, level_ptr;
  IntVec slot_row, rp;  // slot s reads row slot_row[s]; its entries at [rp[s], rp[s + 1])
  IntVec nat_rp;
  std::vector<char, NoInit<char>> store;
  int *nat_ci = nullptr;
  double *nat_cv = nullptr;
  std::shared_ptr<const DblVec> nat_diag;  // pivots, natural order
  std::vector<unsigned short, NoInit<unsigned short>> proxy;  // per slot (slot_proxies)
  int nlevels = 0;
};

// Room for m entries (int index + double value) in one allocation.
void alloc_entries(HostSweep &h, size_t m) {
  const size_t ci_bytes = (m * sizeof(int) + 7) & ~(size_t)7;
  h.store.resize(std::max<size_t>(ci_bytes + m * sizeof(double), 8));
  h.nat_ci = reinterpret_cast<int *>(h.store.data());
  h.nat_cv = reinterpret_cast<double *>(h.store.data() + ci_bytes);
}

// Host threads for the factor analysis: hardware threads, at most
// OMP_NUM_THREADS when that is set, at most 32.
int host_threads() {
  int t = (int)std::thread::hardware_concurrency();
  if (const char *e = getenv("OMP_NUM_THREADS")) {
    const int v = atoi(e);
    if (v > 0) t = std::min(t, v);
  }
  return std::max(1, std::min(t, 32));
}

// Host wall clock in milliseconds (setup timing).
double host_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// f(t) for t in [0, T): T-1 new threads, the caller's thread runs t = 0.
template <typename F>
void par_run(int T, F &&f) {
  if (T <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T - 1);
  for (int t = 1; t < T; t++) th.emplace_back([&f, t]() { f(t); });
  f(0);
  for (auto &x : th) x.join();
}

// f(lo, hi) over [0, n) in contiguous chunks, one per thread (chunks of at
// least `grain` items; the caller's thread takes the first).
template <typename F>
void par_for(size_t n, F &&f, size_t grain = 1 << 16) {
  const int T = (int)std::min<size_t>((size_t)host_threads(), std::max<size_t>(1, n / grain));
  par_run(T, [&](int t) { f(n * t / T, n * (t + 1) / T); });
}

// out[0] = 0, out[i + 1] = out[i] + cnt(i) for i in [0, n): per-thread sums
// of contiguous ranges, then each range's running sum from its offset.
template <typename C>
void par_prefix(size_t n, C &&cnt, IntVec &out) {
  out.resize(n + 1);
  const int T = (int)std::min<size_t>((size_t)host_threads(), std::max<size_t>(1, n / (1 << 16)));
  std::vector<long long> part(T + 1, 0);
  par_run(T, [&](int t) {
    long long s = 0;
    for (size_t i = n * t / T; i < n * (t + 1) / T; i++) s += cnt(i);
    part[t + 1] = s;
  });
  for (int t = 0; t < T; t++) part[t + 1] += part[t];
  out[0] = 0;
  par_run(T, [&](int t) {
    long long s = part[t];
    for (size_t i = n * t / T; i < n * (t + 1) / T; i++) {
      s += cnt(i);
      out[i + 1] = (int)s;
    }
  });
}

// Column ranges for T threads with about equal work (entries + 2 per column).
std::vector<size_t> column_ranges(size_t n, const size_t *colPtr, int T) {
  std::vector<size_t> b(T + 1, n);
  b[0] = 0;
  const double total = (double)colPtr[n] + 2.0 * n;
  size_t j = 0;
  for (int t = 1; t < T; t++) {
    const double goal = total * t / T;
    size_t lo = j, hi = n;
    while (lo < hi) {  // first column whose start reaches the goal
      const size_t mid = (lo + hi) / 2;
      if ((double)colPtr[mid] + 2.0 * mid < goal)
        lo = mid + 1;
      else
        hi = mid;
    }
    b[t] = j = lo;
  }
  return b;
}

// Tasks of a dependency graph on a pool of T threads: task v may run once
// every task u with an edge u -> v has finished (succ_ptr/succ: the edges
// out of each task, indeg: edges into it).  Ready tasks wait on a shared
// stack; piece(t, v) runs task v on thread t.
template <typename F>
void run_dag(int V, const std::vector<int> &succ_ptr, const std::vector<int> &succ,
             std::vector<int> &indeg, int T, F &&piece) {
  std::mutex mu;
  std::vector<int> ready;
  for (int v = 0; v < V; v++)
    if (indeg[v] == 0) ready.push_back(v);
  std::atomic<int> left{V};
  par_run(T, [&](int t) {
    for (;;) {
      int v = -1;
      {
        std::lock_guard<std::mutex> g(mu);
        if (!ready.empty()) {
          v = ready.back();
          ready.pop_back();
        }
      }
      if (v < 0) {
        if (left.load(std::memory_order_acquire) == 0) return;
        std::this_thread::yield();
        continue;
      }
      piece(t, v);
      for (int k = succ_ptr[v]; k < succ_ptr[v + 1]; k++) {
        const int w = succ[k];
        if (__atomic_sub_fetch(&indeg[w], 1, __ATOMIC_ACQ_REL) == 0) {
          std::lock_guard<std::mutex> g(mu);
          ready.push_back(w);
        }
      }
      left.fetch_sub(1, std::memory_order_acq_rel);
    }
  });
}

// Rows sorted by level (stable counting sort: inside a level the rows keep
// their nested-dissection order, so a run of slots is a compact piece of
// the domain), then level-chunked 64-slot tasks.  Threads count their row
// ranges per level when the per-thread counters are no larger than the rows.
void build_slots(const IntVec &lev, int nlev, HostSweep &hs) {
  const int n = (int)lev.size();
  int T = (int)std::min<size_t>((size_t)host_threads(), std::max<size_t>(1, (size_t)n / (1 << 16)));
  if ((long long)T * (nlev + 1) > n) T = 1;
  std::vector<int> cnt((size_t)T * (nlev + 1), 0);
  par_run(T, [&](int t) {
    int *c = &cnt[(size_t)t * (nlev + 1)];
    for (int i = (int)((long long)n * t / T); i < (int)((long long)n * (t + 1) / T); i++) c[lev[i]]++;
  });
  std::vector<int> lp(nlev + 1, 0);
  {  // level l of thread t starts after levels < l and after threads < t in level l
    int s = 0;
    for (int l = 0; l < nlev; l++) {
      lp[l] = s;
      for (int t = 0; t < T; t++) {
        int &c = cnt[(size_t)t * (nlev + 1) + l];
        const int k = c;
        c = s;
        s += k;
      }
    }
    lp[nlev] = s;
  }
  hs.slot_row.resize(n);
  par_run(T, [&](int t) {
    int *c = &cnt[(size_t)t * (nlev + 1)];
    for (int i = (int)((long long)n * t / T); i < (int)((long long)n * (t + 1) / T); i++)
      hs.slot_row[c[lev[i]]++] = i;
  });
  hs.task_ptr.clear();
  for (int l = 0; l < nlev; l++) {
    for (int s = lp[l]; s < lp[l + 1]; s += kTaskRows) hs.task_ptr.push_back(s);
  }
  hs.task_ptr.push_back(n);
  hs.nlevels = nlev;
  hs.level_ptr = std::move(lp);
}

// Inside every row, entries ordered by the level of the row they read,
// latest last, ties by column, then by value.  Single-column sweeps read a
// row from its end, so the lane that polls for a wave watches the value
// produced last and, when it lands, the rest of the row is usually there
// too.  Multi-column sweeps read from the start, several rounds per row: the
// rounds over long-finished entries run while the latest are still in
// flight.  The key fixes the order whatever order the entries arrive in.
struct LevelEnt {
  unsigned long long key;  // level << 32 | column
  double v;
  bool operator<(const LevelEnt &o) const {
    if (key != o.key) return key < o.key;
    unsigned long long a, b;
    memcpy(&a, &v, 8);
    memcpy(&b, &o.v, 8);
    return a < b;
  }
};

void sort_level_ents(LevelEnt *b, int k) {
  // insertion sort: linear on the nearly sorted rows the analysis produces;
  // std::sort takes over once it has moved more than 8 entries per entry
  long long budget = 8LL * k + 64;
  for (int t = 1; t < k; t++) {
    if (!(b[t] < b[t - 1])) continue;
    const LevelEnt x = b[t];
    int u = t;
    for (; u > 0 && x < b[u - 1]; u--) b[u] = b[u - 1];
    b[u] = x;
    budget -= t - u;
    if (budget < 0) {
      std::sort(b, b + k);
      return;
    }
  }
}

// The proxy of a row, for the waits of the persistent sweep.  A row of
// level m stores its k entries latest level last (level_at(u) of entry u
// ascending), so its trailing entries are those of level m - 1; the proxy
// is the distance from the row's end of the entry just before them, when
// that entry lies at most RCHOL_AMD_PROXY_GAP levels below the row's own.
// Once that value is published, the levels just below the row's are being
// computed, and the wave can afford to poll each pending value itself.
// 0: no such entry, a distance beyond 16 bits, or level 0 or 1.  Gap 3
// measured best at 256^3 (forward sweep 1.58 / 1.55 / 1.59 ms for gap
// 2 / 3 / any).
#ifndef RCHOL_AMD_PROXY_GAP
#define RCHOL_AMD_PROXY_GAP 3
#endif
using ProxyVec = std::vector<unsigned short, NoInit<unsigned short>>;

template <typename LV>
inline unsigned short row_proxy(int m, int k, LV &&level_at) {
  if (m < 2) return 0;
  int c = 0;
  while (c < k && level_at(k - 1 - c) == m - 1) c++;
  if (c >= k || c > 0xffff || level_at(k - 1 - c) < m - RCHOL_AMD_PROXY_GAP) return 0;
  return (unsigned short)c;
}

// Per-slot proxies from per-row ones; 0 in levels of kProxyMaxRows rows or
// more (those become stream tasks, which do not use proxies).
constexpr int kProxyMaxRows = 1 << 18;

void slot_proxies(size_t n, const ProxyVec &px_row, HostSweep &h) {
  h.proxy.resize(n);
  const std::vector<int> &lp = h.level_ptr;
  par_for(n, [&](size_t s0, size_t s1) {
    if (s0 >= s1) return;
    int l = (int)(std::upper_bound(lp.begin(), lp.end(), (int)s0) - lp.begin()) - 1;
    for (size_t s = s0; s < s1;) {
      while ((size_t)lp[l + 1] <= s) l++;
      const size_t e = std::min(s1, (size_t)lp[l + 1]);
      if (lp[l + 1] - lp[l] >= kProxyMaxRows)
        std::fill(h.proxy.begin() + s, h.proxy.begin() + e, (unsigned short)0);
      else
        for (size_t q = s; q < e; q++) h.proxy[q] = px_row[h.slot_row[q]];
      s = e;
    }
  });
}

// Host analysis of the stored factor: validation (lower triangle, one
// nonzero diagonal per column), dependency levels of both sweeps, their
// slots, and their rows in natural order with the entries in sweep order.
// Forward: row i of G waits for every column j < i with
// G(i,j) != 0.  Backward (G^T z = y): row i waits for every j > i with
// G(j,i) != 0, i.e. the stored column i.
struct StagedEnt {
  int i, j;
  double v;
};

static int analyze_factor(size_t n, const size_t *colPtr, const size_t *rowIdx, const double *val,
                          HostSweep &hf, HostSweep &hb, int &nlf, int &nlb) {
  const size_t nnz = colPtr[n];
  const size_t noff = nnz >= n ? nnz - n : 0;
  const int T = (int)std::min<size_t>((size_t)host_threads(), std::max<size_t>(1, n / 4096));
  const std::vector<size_t> cr = column_ranges(n, colPtr, T);
  // buckets of 2^bs rows for the transposition: 1024 rows (at most 65536
  // buckets), so that a bucket's rows stay in a core's L2 while placed
  int bs = 10;
  while (((size_t)65536 << bs) < n) bs++;
  const int NB = (int)((n + ((size_t)1 << bs) - 1) >> bs);

  // 1. columns: validation, pivots, entries per (thread, row bucket)
  auto diag_p = std::make_shared<DblVec>(n);
  DblVec &diag_v = *diag_p;
  std::vector<size_t> bcnt((size_t)T * NB, 0);
  size_t bad = n;
  auto column_ok = [&](size_t j) -> bool {
    if (colPtr[j + 1] < colPtr[j] || colPtr[j + 1] > nnz) return false;
    bool seen = false;
    for (size_t e = colPtr[j]; e < colPtr[j + 1]; e++) {
      const size_t i = rowIdx[e];
      if (i >= n || i < j || (i == j && seen)) return false;
      if (i == j) {
        if (val[e] == 0.0) return false;
        seen = true;
      }
    }
    return seen;
  };
  par_run(T, [&](int t) {
    size_t *bc = &bcnt[(size_t)t * NB];
    for (size_t j = cr[t]; j < cr[t + 1]; j++) {
      if (!column_ok(j)) {  // the lowest bad column of this range
        size_t cur = __atomic_load_n(&bad, __ATOMIC_RELAXED);
        while (j < cur &&
               !__atomic_compare_exchange_n(&bad, &cur, j, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        return;
      }
      for (size_t e = colPtr[j]; e < colPtr[j + 1]; e++) {
        const size_t i = rowIdx[e];
        if (i == j) {
          diag_v[j] = val[e];
          continue;
        }
        bc[i >> bs]++;
      }
    }
  });
  if (bad < n) {  // the error names the lowest bad column, as a sequential pass would
    const size_t j = bad;
    if (colPtr[j + 1] < colPtr[j]) return fail(RCHOL_AMD_EINVAL, "colPtr not monotone at %zu", j);
    if (colPtr[j + 1] > nnz)
      return fail(RCHOL_AMD_EINVAL, "colPtr[%zu] = %zu exceeds nnz = %zu", j + 1,
                  (size_t)colPtr[j + 1], nnz);
    bool seen = false;
    for (size_t e = colPtr[j]; e < colPtr[j + 1]; e++) {
      const size_t i = rowIdx[e];
      if (i >= n || i < j)
        return fail(RCHOL_AMD_EINVAL, "G entry (%zu,%zu) is not in the lower triangle", i, j);
      if (i == j) {
        if (seen) return fail(RCHOL_AMD_EINVAL, "column %zu holds two diagonal entries", j);
        seen = true;
      }
    }
    return fail(RCHOL_AMD_EPIVOT, "G(%zu,%zu) is zero or missing", j, j);
  }

  // 2. CSR of lower G in natural row order: entries staged by row bucket
  // (bucket-major, then thread, then column order), then each bucket placed
  // by row.  Inside a row the columns ascend.  The staging memory is the
  // backward sweep's store (its entries are written in step 3).
  std::vector<size_t> bst(NB + 1, 0);  // staged range of each bucket
  std::vector<size_t> cur((size_t)T * NB);
  {
    size_t s = 0;
    for (int b = 0; b < NB; b++) {
      bst[b] = s;
      for (int t = 0; t < T; t++) {
        cur[(size_t)t * NB + b] = s;
        s += bcnt[(size_t)t * NB + b];
      }
    }
    bst[NB] = s;
  }
  hf.nat_rp.resize(n + 1);
  alloc_entries(hf, noff);
  hb.store.resize(std::max<size_t>(noff * sizeof(StagedEnt), 8));
  {
    int *cin = hf.nat_ci;
    double *cvn = hf.nat_cv;
    int *rpn = hf.nat_rp.data();
    StagedEnt *st = reinterpret_cast<StagedEnt *>(hb.store.data());
    par_run(T, [&](int t) {
      size_t *c = &cur[(size_t)t * NB];
      for (size_t j = cr[t]; j < cr[t + 1]; j++) {
        for (size_t e = colPtr[j]; e < colPtr[j + 1]; e++) {
          const size_t i = rowIdx[e];
          if (i == j) continue;
          st[c[i >> bs]++] = {(int)i, (int)j, val[e]};
        }
      }
    });
    std::atomic<int> nextb{0};
    par_run(std::min(T, NB), [&](int) {
      std::vector<int> cnt((size_t)1 << bs);
      for (int b; (b = nextb.fetch_add(1, std::memory_order_relaxed)) < NB;) {
        const size_t r0 = (size_t)b << bs, r1 = std::min(n, r0 + ((size_t)1 << bs));
        std::fill(cnt.begin(), cnt.begin() + (r1 - r0), 0);
        for (size_t k = bst[b]; k < bst[b + 1]; k++) cnt[st[k].i - r0]++;
        size_t s = bst[b];
        for (size_t r = r0; r < r1; r++) {
          rpn[r] = (int)s;
          const int c = cnt[r - r0];
          cnt[r - r0] = (int)s;
          s += c;
        }
        for (size_t k = bst[b]; k < bst[b + 1]; k++) {
          const int pos = cnt[st[k].i - r0]++;
          cin[pos] = st[k].j;
          cvn[pos] = st[k].v;
        }
      }
    });
    rpn[n] = (int)noff;
  }
  // the backward sweep's rows are the stored columns: column j's
  // off-diagonal entries at colPtr[j] - j
  static_assert(sizeof(StagedEnt) >= sizeof(int) + sizeof(double), "staging too small");
  {
    const size_t ci_bytes = (noff * sizeof(int) + 7) & ~(size_t)7;
    hb.nat_ci = reinterpret_cast<int *>(hb.store.data());
    hb.nat_cv = reinterpret_cast<double *>(hb.store.data() + ci_bytes);
    hb.nat_rp.resize(n + 1);
    par_for(n + 1, [&](size_t j0, size_t j1) {
      for (size_t j = j0; j < j1; j++) hb.nat_rp[j] = (int)(colPtr[j] - j);
    });
  }

  // Pieces for the level passes: runs of rows of about equal work (both
  // sweeps' entries of a row, plus 2), about 64 per thread, so a dense
  // separator spreads over several pieces.  A piece also ends where a run
  // of rows that read nothing begins: a nested-dissection block starts so
  // (its first independent set, or its first row in the given order), and a
  // piece across two blocks would chain them.  A bitmap per thread marks
  // which pieces of columns feed which pieces of rows.
  constexpr int kMaxPieces = 4096;
  IntVec piece_of;
  std::vector<size_t> pstart;
  int P = 0;
  std::vector<unsigned long long> bm;
  size_t words = 0;
  {
    const IntVec &nr = hf.nat_rp;
    auto weight = [&](size_t r) {
      return (size_t)(nr[r + 1] - nr[r]) + (size_t)(colPtr[r + 1] - colPtr[r] - 1) + 2;
    };
    auto reads = [&](size_t r) { return nr[r + 1] > nr[r]; };
    std::vector<size_t> sum(T + 1, 0);
    par_run(T, [&](int t) {
      size_t w = 0;
      for (size_t r = n * t / T; r < n * (t + 1) / T; r++) w += weight(r);
      sum[t + 1] = w;
    });
    for (int t = 0; t < T; t++) sum[t + 1] += sum[t];
    const size_t target = std::max<size_t>(1, sum[T] / (64 * (size_t)T));
    IntVec cut(n), first;  // first[r] = cuts in rows [0, r)
    for (int blocks = 1; blocks >= 0; blocks--) {
      par_run(T, [&](int t) {
        size_t w = sum[t];  // work of the rows before r
        for (size_t r = n * t / T; r < n * (t + 1) / T; r++) {
          bool c = r == 0;
          if (!c) {
            const size_t wp = w - weight(r - 1);
            c = w / target != wp / target || (blocks && !reads(r) && reads(r - 1));
          }
          cut[r] = c ? 1 : 0;
          w += weight(r);
        }
      });
      par_prefix(n, [&](size_t r) { return cut[r]; }, first);
      P = first[n];
      if (P <= kMaxPieces) break;  // else too many block starts: by work only
    }
    pstart.assign(P + 1, n);
    piece_of.resize(n);
    par_for(n, [&](size_t r0, size_t r1) {
      for (size_t r = r0; r < r1; r++) {
        piece_of[r] = first[r + 1] - 1;
        if (cut[r]) pstart[first[r]] = r;
      }
    });
    words = ((size_t)P * P + 63) / 64;
    bm.assign((size_t)T * words, 0);
    par_run(T, [&](int t) {
      unsigned long long *w = &bm[(size_t)t * words];
      for (size_t i = n * t / T; i < n * (t + 1) / T; i++) {
        const size_t pi = (size_t)piece_of[i];
        for (int e = nr[i]; e < nr[i + 1]; e++) {
          const size_t pj = (size_t)piece_of[hf.nat_ci[e]];
          if (pj == pi) continue;
          const size_t bit = pj * P + pi;
          w[bit >> 6] |= 1ull << (bit & 63);
        }
      }
    });
  }

  // 3. levels, and each row's entries sorted by level.  Forward: lev(i) =
  // 1 + max lev(j) over row i (0 if empty), rows ascending, sorted in a pass
  // of their own afterwards; backward: lev(j) = 1 + max lev(i) over column
  // j, columns descending, each sorted as soon as its level is known.  Task q < P is
  // forward piece q, task P + q backward piece q; an entry from column
  // piece a to row piece b > a makes forward b wait for forward a and
  // backward a wait for backward b.
  IntVec levf(n), levb(n);
  ProxyVec pxf(n), pxb(n);  // per-row proxies (row_proxy)
  std::vector<int> topf(T, 0), topb(T, 0);
  {
    for (int t = 1; t < T; t++)
      for (size_t k = 0; k < words; k++) bm[k] |= bm[(size_t)t * words + k];
    std::vector<int> deg_out(2 * P + 1, 0), indeg(2 * P, 0);
    auto each_edge = [&](auto &&f) {
      for (int a = 0; a < P; a++)
        for (int b = a + 1; b < P; b++) {
          const size_t bit = (size_t)a * P + b;
          if (bm[bit >> 6] >> (bit & 63) & 1) f(a, b);
        }
    };
    each_edge([&](int a, int b) {
      deg_out[a + 1]++;      // forward a -> b
      deg_out[P + b + 1]++;  // backward b -> a
      indeg[b]++;
      indeg[P + a]++;
    });
    for (int v = 0; v < 2 * P; v++) deg_out[v + 1] += deg_out[v];
    std::vector<int> succ(deg_out[2 * P]), at(deg_out.begin(), deg_out.end() - 1);
    each_edge([&](int a, int b) {
      succ[at[a]++] = b;
      succ[at[P + b]++] = P + a;
    });
    const int *rpn = hf.nat_rp.data();
    int *cin = hf.nat_ci, *bci = hb.nat_ci;
    double *cvn = hf.nat_cv, *bcv = hb.nat_cv;
    run_dag(2 * P, deg_out, succ, indeg, T, [&](int t, int v) {
      int top = 0;
      if (v < P) {
        for (size_t i = pstart[v]; i < pstart[v + 1]; i++) {
          int m = 0;
          for (int e = rpn[i]; e < rpn[i + 1]; e++) m = std::max(m, levf[cin[e]] + 1);
          levf[i] = m;
          top = std::max(top, m);
        }
        topf[t] = std::max(topf[t], top);
      } else {
        // a backward row (stored column) is final once its level is: it is
        // sorted and stored here, where the pool has room beside the
        // forward sweep's chain of separator pieces.  Read from its end:
        // later rows finish earlier in this sweep.
        std::vector<LevelEnt> buf;  // one per task: no cache line shared between threads
        for (size_t j = pstart[v - P + 1]; j-- > pstart[v - P];) {
          buf.clear();
          int m = 0;
          for (size_t e = colPtr[j + 1]; e-- > colPtr[j];) {
            const size_t i = rowIdx[e];
            if (i == j) continue;
            const int l = levb[i];
            m = std::max(m, l + 1);
            buf.push_back({((unsigned long long)(unsigned)l << 32) | (unsigned)i, val[e]});
          }
          levb[j] = m;
          top = std::max(top, m);
          const int k = (int)buf.size();
          sort_level_ents(buf.data(), k);
          pxb[j] = row_proxy(m, k, [&](int u) { return (int)(buf[u].key >> 32); });
          const size_t o = colPtr[j] - j;
          for (int u = 0; u < k; u++) {
            bci[o + u] = (int)(unsigned)(buf[u].key & 0xffffffffu);
            bcv[o + u] = buf[u].v;
          }
        }
        topb[t] = std::max(topb[t], top);
      }
    });
    // Forward rows in sweep order, in place, rows in parallel (chunks taken
    // from a shared counter: the dense separator rows come last); most
    // already are.
    std::atomic<size_t> next{0};
    constexpr size_t kRowChunk = 4096;
    par_run(T, [&](int) {
      std::vector<LevelEnt> buf;
      for (size_t r0; (r0 = next.fetch_add(kRowChunk, std::memory_order_relaxed)) < n;) {
        const size_t r1 = std::min(n, r0 + kRowChunk);
        for (size_t i = r0; i < r1; i++) {
          const int e0 = rpn[i], k = rpn[i + 1] - e0;
          bool sorted = true;
          unsigned long long prev = 0;
          for (int u = 0; u < k && sorted; u++) {
            const int j = cin[e0 + u];
            const unsigned long long key = ((unsigned long long)(unsigned)levf[j] << 32) | (unsigned)j;
            sorted = u == 0 || key > prev;
            prev = key;
          }
          if (!sorted) {
            buf.resize(k);
            for (int u = 0; u < k; u++) {
              const int j = cin[e0 + u];
              buf[u] = {((unsigned long long)(unsigned)levf[j] << 32) | (unsigned)j, cvn[e0 + u]};
            }
            sort_level_ents(buf.data(), k);
            for (int u = 0; u < k; u++) {
              cin[e0 + u] = (int)(unsigned)(buf[u].key & 0xffffffffu);
              cvn[e0 + u] = buf[u].v;
            }
          }
          pxf[i] = row_proxy(levf[i], k, [&](int u) { return levf[cin[e0 + u]]; });
        }
      }
    });
  }
  nlf = *std::max_element(topf.begin(), topf.end()) + 1;
  nlb = *std::max_element(topb.begin(), topb.end()) + 1;

  // 4. slots, each slot's entry range in the slot-order layout, proxies
  build_slots(levf, nlf, hf);
  build_slots(levb, nlb, hb);
  for (HostSweep *h : {&hf, &hb}) {
    const IntVec &nr = h->nat_rp;
    const IntVec &sr = h->slot_row;
    par_prefix(n, [&](size_t s) { return nr[sr[s] + 1] - nr[sr[s]]; }, h->rp);
    h->nat_diag = diag_p;
    slot_proxies(n, h == &hf ? pxf : pxb, *h);
  }
  return RCHOL_AMD_OK;
}
