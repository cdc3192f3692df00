// Host analysis of a stored factor for the triangular sweeps: dependency
// levels of the forward and the backward sweep, rows in level ("slot")
// order and their off-diagonal entries in that order.  Plain C++ (no HIP):
// included by engine.hip inside its anonymous namespace, after fail() and
// kTaskRows.
#pragma once

// Host-side schedule of one sweep before upload.
// Allocator whose resize leaves new elements uninitialised: the analysis
// writes every element of the entry arrays from several threads, and a
// zero-fill first would be a serial pass over gigabytes.
template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <typename U>
  NoInit(const NoInit<U> &) {}
  template <typename U>
  void construct(U *p) noexcept {
    ::new ((void *)p) U;
  }
  template <typename U, typename... Args>
  void construct(U *p, Args &&...args) {
    ::new ((void *)p) U(std::forward<Args>(args)...);
  }
};

struct HostSweep {
  std::vector<int> task_ptr, slot_row, rp, level_ptr;
  std::vector<int, NoInit<int>> ci;
  std::vector<double, NoInit<double>> cv;
  std::vector<double> diag;
  int nlevels = 0;
};

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

// f(lo, hi) over [0, n) in contiguous chunks, one per thread (chunks of at
// least `grain` items; the caller's thread takes the first).
template <typename F>
void par_for(size_t n, F &&f, size_t grain = 1 << 16) {
  const size_t T = std::min<size_t>((size_t)host_threads(), std::max<size_t>(1, n / grain));
  if (T <= 1) {
    f((size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T - 1);
  for (size_t t = 1; t < T; t++) th.emplace_back([&f, n, t, T]() { f(n * t / T, n * (t + 1) / T); });
  f((size_t)0, n / T);
  for (auto &x : th) x.join();
}

// Rows sorted by level (stable counting sort: inside a level the rows keep
// their nested-dissection order, so a run of slots is a compact piece of
// the domain), then level-chunked 64-slot tasks.
void build_slots(const std::vector<int> &lev, int nlev, HostSweep &hs, std::vector<int> &slot_of) {
  const int n = (int)lev.size();
  std::vector<int> cnt(nlev + 1, 0);
  for (int i = 0; i < n; i++) cnt[lev[i] + 1]++;
  for (int l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
  hs.slot_row.assign(n, 0);
  slot_of.assign(n, 0);
  std::vector<int> cur(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; i++) {
    int s = cur[lev[i]]++;
    hs.slot_row[s] = i;
    slot_of[i] = s;
  }
  hs.task_ptr.clear();
  for (int l = 0; l < nlev; l++) {
    for (int s = cnt[l]; s < cnt[l + 1]; s += kTaskRows) hs.task_ptr.push_back(s);
  }
  hs.task_ptr.push_back(n);
  hs.nlevels = nlev;
  hs.level_ptr.assign(cnt.begin(), cnt.end());
}

// Inside every row, entries ordered by the level of the row they read,
// latest last, ties by column, then by position (rows in parallel; short
// rows by insertion sort, long ones by std::sort on local keys).
// Single-column sweeps read a row from its end, so the lane that polls for
// a wave watches the value produced last and, when it lands, the rest of
// the row is usually there too.  Multi-column sweeps read from the start,
// several rounds per row: the rounds over long-finished entries run while
// the latest are still in flight.  The key fixes the order whatever order
// the entries arrive in.
void order_entries_by_level(HostSweep &hs, const std::vector<int> &lev) {
  struct Ent {
    unsigned long long key;  // level << 32 | column
    unsigned pos;
    double v;
    bool operator<(const Ent &o) const { return key != o.key ? key < o.key : pos < o.pos; }
  };
  const size_t n = hs.slot_row.size();
  par_for(n, [&](size_t s0, size_t s1) {
    std::vector<Ent> buf;
    for (size_t s = s0; s < s1; s++) {
      const int e0 = hs.rp[s], e1 = hs.rp[s + 1];
      const int k = e1 - e0;
      if (k < 2) continue;
      buf.resize(k);
      bool sorted = true;
      for (int t = 0; t < k; t++) {
        const int j = hs.ci[e0 + t];
        buf[t] = {((unsigned long long)(unsigned)lev[j] << 32) | (unsigned)j, (unsigned)t,
                  hs.cv[e0 + t]};
        if (t && buf[t] < buf[t - 1]) sorted = false;
      }
      if (sorted) continue;
      if (k <= 16) {
        for (int t = 1; t < k; t++) {
          const Ent x = buf[t];
          int u = t;
          for (; u > 0 && x < buf[u - 1]; u--) buf[u] = buf[u - 1];
          buf[u] = x;
        }
      } else {
        std::sort(buf.begin(), buf.end());
      }
      for (int t = 0; t < k; t++) {
        hs.ci[e0 + t] = (int)(unsigned)(buf[t].key & 0xffffffffu);
        hs.cv[e0 + t] = buf[t].v;
      }
    }
  });
}

// Host analysis of the stored factor: validation (lower triangle, one
// nonzero diagonal per column), dependency levels of both sweeps and their
// slot-ordered CSR.  Forward: row i of G waits for every column j < i with
// G(i,j) != 0.  Backward (G^T z = y): row i waits for every j > i with
// G(j,i) != 0, i.e. the stored column i.  Host threads split the columns or
// slots; the two level passes (sequential in the column order) run side by
// side; the result does not depend on the thread count.
static int analyze_factor(size_t n, const size_t *colPtr, const size_t *rowIdx, const double *val,
                          HostSweep &hf, HostSweep &hb, int &nlf, int &nlb) {
  const size_t nnz = colPtr[n];
  // Validate lower-triangular columns and find the diagonal of each; the
  // error names the lowest bad column, as a sequential pass would.
  std::vector<double> diag(n, 0.0);
  std::vector<int> offf(n, 0);  // off-diagonals per row of G (forward sweep)
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
  auto report_column = [&](size_t j) -> int {
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
  };
  {
    size_t bad = n;
    par_for(n, [&](size_t j0, size_t j1) {
      for (size_t j = j0; j < j1; j++) {
        if (column_ok(j)) continue;
        size_t cur = __atomic_load_n(&bad, __ATOMIC_RELAXED);
        while (j < cur &&
               !__atomic_compare_exchange_n(&bad, &cur, j, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        break;
      }
    });
    if (bad < n) return report_column(bad);
    par_for(n, [&](size_t j0, size_t j1) {
      for (size_t j = j0; j < j1; j++) {
        for (size_t e = colPtr[j]; e < colPtr[j + 1]; e++) {
          const size_t i = rowIdx[e];
          if (i == j)
            diag[j] = val[e];
          else
            __atomic_fetch_add(&offf[i], 1, __ATOMIC_RELAXED);
        }
      }
    });
  }
  std::vector<int> levf(n, 0), levb(n, 0);
  nlf = 0;
  nlb = 0;
  std::thread fwd_levels([&]() {
    int m = 0;
    for (size_t j = 0; j < n; j++) {
      const int lj = levf[j] + 1;
      m = std::max(m, lj);
      for (size_t e = colPtr[j]; e < colPtr[j + 1]; e++) {
        size_t i = rowIdx[e];
        if (i != j && levf[i] < lj) levf[i] = lj;
      }
    }
    nlf = m;
  });
  for (size_t jj = n; jj-- > 0;) {
    int m = 0;
    for (size_t e = colPtr[jj]; e < colPtr[jj + 1]; e++) {
      size_t i = rowIdx[e];
      if (i != jj) m = std::max(m, levb[i] + 1);
    }
    levb[jj] = m;
    nlb = std::max(nlb, m + 1);
  }
  fwd_levels.join();
  const size_t noff = nnz - n;
  std::vector<int> slot_of_f, slot_of_b;
  {
    std::thread fwd_slots([&]() { build_slots(levf, nlf, hf, slot_of_f); });
    build_slots(levb, nlb, hb, slot_of_b);
    fwd_slots.join();
  }
  // forward sweep: CSR of lower G in slot order.  A thread owns a range of
  // rows and scans the columns that can reach it (j < its last row), so
  // every row is written by one thread, in column order.
  hf.rp.assign(n + 1, 0);
  for (size_t s = 0; s < n; s++) hf.rp[s + 1] = hf.rp[s] + offf[hf.slot_row[s]];
  hf.ci.resize(noff);
  hf.cv.resize(noff);
  hf.diag.assign(n, 0.0);
  {
    std::vector<int> cur(n);
    par_for(n, [&](size_t i0, size_t i1) {
      for (size_t i = i0; i < i1; i++) cur[i] = hf.rp[slot_of_f[i]];
      for (size_t j = 0; j < i1; j++) {
        for (size_t e = colPtr[j]; e < colPtr[j + 1]; e++) {
          const size_t i = rowIdx[e];
          if (i < i0 || i >= i1 || i == j) continue;
          const int pos = cur[i]++;
          hf.ci[pos] = (int)j;
          hf.cv[pos] = val[e];
        }
      }
    });
    par_for(n, [&](size_t s0, size_t s1) {
      for (size_t s = s0; s < s1; s++) hf.diag[s] = diag[hf.slot_row[s]];
    });
  }
  // backward sweep: the stored rows (CSR of G^T) in slot order
  hb.rp.assign(n + 1, 0);
  for (size_t s = 0; s < n; s++) {
    int row = hb.slot_row[s];
    hb.rp[s + 1] = hb.rp[s] + (int)(colPtr[row + 1] - colPtr[row] - 1);
  }
  hb.ci.resize(noff);
  hb.cv.resize(noff);
  hb.diag.assign(n, 0.0);
  par_for(n, [&](size_t s0, size_t s1) {
    for (size_t s = s0; s < s1; s++) {
      int row = hb.slot_row[s];
      int pos = hb.rp[s];
      for (size_t e = colPtr[row]; e < colPtr[row + 1]; e++) {
        size_t j = rowIdx[e];
        if (j == (size_t)row) continue;
        hb.ci[pos] = (int)j;
        hb.cv[pos] = val[e];
        pos++;
      }
      hb.diag[s] = diag[row];
    }
  });
  order_entries_by_level(hf, levf);
  order_entries_by_level(hb, levb);
  return RCHOL_AMD_OK;
}

