// Host plan of the chunked sweeps (k_sptrsv_chunk).  Included by engine.hip
// inside its anonymous namespace, after HostSweep and par_for.
//
// Input: both sweeps' rows with their off-diagonal entries in natural row
// order (HostSweep: row i's entries at [nat_rp[i], nat_rp[i + 1]) of
// nat_ci / nat_cv, in sweep order) and pivots.  The forward sweep reads
// entries j < i, the backward sweep j > i.

// Bounds of the chunks: one partition of [0, n) into index ranges shared by
// both sweeps.  A range ends at S rows, and, once it holds kChunkMinRows,
// where the next rows start a new nested-dissection block: either the next
// kChunkWindow rows read nothing before that point (a block that depends on
// nothing earlier: a leaf), or a row reaches before the range while the row
// before it did not (a separator after its children).  Cuts elsewhere only
// change the speed, never the result.
constexpr int kChunkMinRows = 256;

// f(i) for i in [0, n) on the host threads, indices handed out one at a time
// (chunks differ in size by orders of magnitude).
template <typename F>
void par_each(size_t n, F &&f) {
  const size_t T = std::min<size_t>((size_t)host_threads(), n);
  std::atomic<size_t> next{0};
  auto run = [&]() {
    for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
  };
  std::vector<std::thread> th;
  for (size_t t = 1; t < T; t++) th.emplace_back(run);
  run();
  for (auto &x : th) x.join();
}
constexpr int kChunkWindow = 64;

std::vector<int> chunk_bounds(const HostSweep &hf, int n, int S) {
  std::vector<int> lo(n);
  par_for((size_t)n, [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; i++) {
      int m = (int)i;
      for (int e = hf.nat_rp[i]; e < hf.nat_rp[i + 1]; e++) m = std::min(m, hf.nat_ci[e]);
      lo[i] = m;
    }
  });
  // wmin[e] = min(lo[e .. e + kChunkWindow)): sliding minimum, right to left
  std::vector<int> wmin(n), dq(n);
  {
    int head = 0, tail = 0;  // dq[head, tail): indices, decreasing, lo increasing
    for (int e = n - 1; e >= 0; e--) {
      while (tail > head && lo[dq[tail - 1]] >= lo[e]) tail--;
      dq[tail++] = e;
      while (dq[head] >= e + kChunkWindow) head++;
      wmin[e] = lo[dq[head]];
    }
  }
  std::vector<int> b{0};
  int c0 = 0;
  for (int e = 1; e < n; e++) {
    const int sz = e - c0;
    bool cut = sz >= S;
    if (!cut && sz >= kChunkMinRows)
      cut = wmin[e] >= e || (lo[e] < c0 && lo[e - 1] >= c0);
    if (cut) {
      b.push_back(e);
      c0 = e;
    }
  }
  b.push_back(n);
  return b;
}

// External entries per part (see plan_chunks); a chunk with more than two
// parts' worth hands its external sums to parts.
constexpr long long kExtPartEntries = 8192;

struct HostChunks {
  int nchunks = 0, nparts = 0;
  std::vector<int> c_part0, p_chunk, p_l0, p_l1, p_lg;  // parts of the external sums
  IntVec x_rp;  // per chunk-local row (q): its first external entry
  std::vector<int> order, q0, m, r0, r1, split;
  std::vector<int> dep_ptr, dep;  // the work items each work item waits for (CSR)
  IntVec q_row;
  DblVec q_diag;
  std::vector<int, NoInit<int>> x_l, x_ci;
  std::vector<double, NoInit<double>> x_cv;
  std::vector<unsigned, NoInit<unsigned>> r_w;
  std::vector<double, NoInit<double>> r_v;
  long long rounds = 0;
  double est_us = 0.0;  // critical path estimate of one sweep
};

// Smallest round >= r with a free lane (union-find over full rounds).
struct RoundFill {
  std::vector<int> fill, next;
  int find(int r) {
    grow(r);
    int a = r;
    while (next[a] != a) {
      a = next[a];
      grow(a);
    }
    while (next[r] != a) {  // path compression
      const int t = next[r];
      next[r] = a;
      r = t;
    }
    return a;
  }
  void grow(int r) {
    while ((int)fill.size() <= r) {
      next.push_back((int)fill.size());
      fill.push_back(0);
    }
  }
  int take(int need) {
    const int r = find(need);
    if (++fill[r] == 64) {
      grow(r + 1);
      next[r] = r + 1;
    }
    return r;
  }
};

// Cost model (one sweep, microseconds): a chunk's fixed cost, a round of the
// chain wave, a batch of external gathers per thread.  Measured orders of
// magnitude on MI355X; only the choice between the two sweep forms uses it.
constexpr double kChunkFixedUs = 4.0;
constexpr double kChunkRoundUs = 0.08;
constexpr double kChunkGatherUs = 2.0;

// Plan of one sweep over the bounds b.  backward: the rows read j > i and
// the chunks run from the last range down.
void plan_chunks(const HostSweep &h, int n, const std::vector<int> &b, bool backward,
                 HostChunks &out) {
  const int C = (int)b.size() - 1;
  out.nchunks = C;
  out.q0.assign(C, 0);
  out.m.assign(C, 0);
  out.r0.assign(C, 0);
  out.r1.assign(C, 0);
  IntVec lidx(n);  // every row's entry is written in pass 1
  std::vector<int> rounds_of(C, 0), ext_of(C, 0);
  const int *hci = h.nat_ci;
  const double *hcv = h.nat_cv;
  const DblVec &hdiag = *h.nat_diag;
  std::vector<int, NoInit<int>> ent_round((size_t)h.nat_rp[n]);
  auto row_entries = [&](int i, int &e0, int &e1) {
    e0 = h.nat_rp[i];
    e1 = h.nat_rp[i + 1];
  };
  // pass 1: local order, local index, rounds of each internal entry
  par_each((size_t)C, [&](size_t c) {
    thread_local std::vector<int> llev, cnt, fin, rows;
    {
      const int a = b[c], z = b[c + 1], m = z - a;
      llev.assign(m, 0);
      int maxlev = 0, ext = 0;
      for (int k = 0; k < m; k++) {
        const int i = backward ? z - 1 - k : a + k;
        int e0, e1, lev = 0;
        row_entries(i, e0, e1);
        for (int e = e0; e < e1; e++) {
          const int j = hci[e];
          if (j >= a && j < z)
            lev = std::max(lev, llev[j - a] + 1);
          else
            ext++;
        }
        llev[i - a] = lev;
        maxlev = std::max(maxlev, lev);
      }
      cnt.assign(maxlev + 2, 0);
      for (int k = 0; k < m; k++) cnt[llev[k] + 1]++;
      for (int L = 0; L <= maxlev; L++) cnt[L + 1] += cnt[L];
      rows.assign(m, 0);
      for (int k = 0; k < m; k++) {
        const int i = backward ? z - 1 - k : a + k;
        const int l = cnt[llev[i - a]]++;
        lidx[i] = l;
        rows[l] = i;
      }
      RoundFill rf;
      fin.assign(m, -1);
      for (int l = 0; l < m; l++) {
        int e0, e1;
        row_entries(rows[l], e0, e1);
        for (int e = e0; e < e1; e++) {
          const int j = hci[e];
          if (j < a || j >= z) continue;
          const int r = rf.take(fin[lidx[j]] + 1);
          ent_round[e] = r;
          fin[l] = std::max(fin[l], r);
        }
      }
      int R = 0;
      for (int l = 0; l < m; l++) R = std::max(R, fin[l] + 1);
      // whole prefetch groups (the chain wave's ring has no bounds test)
      rounds_of[c] = (R + kChunkPrefetch - 1) / kChunkPrefetch * kChunkPrefetch;
      ext_of[c] = ext;
    }
  });
  // prefix sums in chunk-id order
  std::vector<long long> xbase(C + 1, 0);
  long long rtot = 0;
  for (int c = 0; c < C; c++) {
    out.q0[c] = b[c];
    out.m[c] = b[c + 1] - b[c];
    out.r0[c] = (int)rtot;
    rtot += rounds_of[c];
    out.r1[c] = (int)rtot;
    xbase[c + 1] = xbase[c] + ext_of[c];
  }
  out.rounds = rtot;
  out.q_row.resize(n);  // pass 2 writes every row of every chunk
  out.q_diag.resize(n);
  out.x_rp.resize(n + 1);
  out.x_l.resize(xbase[C]);
  out.x_ci.resize(xbase[C]);
  out.x_cv.resize(xbase[C]);
  // kChunkPrefetch rounds of padding past the last chunk: the ring's reads
  // run that far ahead
  out.r_w.resize((size_t)(rtot + kChunkPrefetch) * 64);
  out.r_v.resize((size_t)(rtot + kChunkPrefetch) * 64);
  for (size_t e = (size_t)rtot * 64; e < out.r_w.size(); e++) {
    out.r_w[e] = kChunkPad;
    out.r_v[e] = 0.0;
  }
  out.split.assign((size_t)C * (kChunkThreads + 1), 0);
  out.x_rp[n] = (int)xbase[C];
  // pass 2: fill
  par_each((size_t)C, [&](size_t c) {
    thread_local std::vector<int> lane, rows;
    {
      const int a = b[c], z = b[c + 1], m = z - a;
      // this chunk's rounds start empty (the arrays are not initialised)
      std::fill(out.r_w.begin() + (size_t)out.r0[c] * 64, out.r_w.begin() + (size_t)out.r1[c] * 64,
                kChunkPad);
      std::fill(out.r_v.begin() + (size_t)out.r0[c] * 64, out.r_v.begin() + (size_t)out.r1[c] * 64,
                0.0);
      rows.assign(m, 0);
      for (int i = a; i < z; i++) rows[lidx[i]] = i;
      lane.assign(rounds_of[c], 0);
      long long xp = xbase[c];
      for (int l = 0; l < m; l++) {
        const int i = rows[l];
        const double d = hdiag[i];
        out.q_row[a + l] = i;
        out.q_diag[a + l] = d;
        out.x_rp[a + l] = (int)xp;
        for (int e = h.nat_rp[i]; e < h.nat_rp[i + 1]; e++) {
          const int j = hci[e];
          if (j < a || j >= z) {
            out.x_l[xp] = l;
            out.x_ci[xp] = j;
            out.x_cv[xp] = hcv[e];
            xp++;
            continue;
          }
          const int r = ent_round[e];
          const size_t slot = ((size_t)out.r0[c] + r) * 64 + lane[r]++;
          out.r_w[slot] = (unsigned)lidx[j] | ((unsigned)l << 16);
          out.r_v[slot] = hcv[e] / d;
        }
      }
      // a row-aligned, balanced split of the external entries over the threads
      const long long x0 = xbase[c], x1 = xbase[c + 1];
      int *sp = &out.split[c * (size_t)(kChunkThreads + 1)];
      long long k = x0;
      for (int t = 0; t <= kChunkThreads; t++) {
        long long target = x0 + (x1 - x0) * t / kChunkThreads;
        if (target < k) target = k;
        while (target < x1 && target > x0 && out.x_l[target] == out.x_l[target - 1]) target++;
        sp[t] = (int)target;
        k = target;
      }
      sp[kChunkThreads] = (int)x1;
    }
  });
  // external parts: a chunk with many external entries hands their sums to
  // parts, separate work items that other workgroups run in parallel, each
  // over a row-aligned range of the chunk's local rows (it writes the sum of
  // every row of its range, zero for rows without entries, with row groups
  // of 2^p_lg lanes); the chunk then starts from those sums
  std::vector<int> part0(C + 1, 0);
  for (int c = 0; c < C; c++) {
    const long long E = xbase[c + 1] - xbase[c];
    part0[c + 1] = part0[c] + (E > 2 * kExtPartEntries
                                   ? (int)((E + kExtPartEntries - 1) / kExtPartEntries)
                                   : 0);
  }
  const int P = part0[C];
  out.nparts = P;
  out.c_part0 = part0;
  out.p_chunk.assign(P, 0);
  out.p_l0.assign(P, 0);
  out.p_l1.assign(P, 0);
  out.p_lg.assign(P, 0);
  // a row-aligned boundary at or after entry k of chunk c
  auto row_aligned = [&](int c, long long k) {
    const long long x0 = xbase[c], x1 = xbase[c + 1];
    while (k < x1 && k > x0 && out.x_l[k] == out.x_l[k - 1]) k++;
    return k;
  };
  par_each((size_t)C, [&](size_t c) {
    {
      const int np = part0[c + 1] - part0[c];
      if (np == 0) return;
      const long long x0 = xbase[c], E = xbase[c + 1] - x0;
      std::vector<long long> xb(np + 1);
      for (int q = 0; q <= np; q++) xb[q] = row_aligned((int)c, x0 + E * q / np);
      xb[np] = xbase[c + 1];
      for (int q = 0; q < np; q++) {
        const int p = part0[c] + q;
        out.p_chunk[p] = (int)c;
        out.p_l0[p] = q == 0 ? 0 : (xb[q] < xbase[c + 1] ? out.x_l[xb[q]] : out.m[c]);
      }
      for (int q = 0; q < np; q++) {
        const int p = part0[c] + q;
        out.p_l1[p] = q + 1 < np ? out.p_l0[p + 1] : out.m[c];
        // lanes per row: the mean row of the part in one round of
        // kChunkExtBatch gathers per lane
        const int rows = std::max(1, out.p_l1[p] - out.p_l0[p]);
        const long long ents = xb[q + 1] - xb[q];
        int lg = 0;
        while (lg < 6 && (long long)(kChunkExtBatch << lg) * rows < ents) lg++;
        out.p_lg[p] = lg;
      }
      // the chunk's own run of external entries is then empty
      int *sp = &out.split[c * (size_t)(kChunkThreads + 1)];
      for (int t = 0; t <= kChunkThreads; t++) sp[t] = (int)x0;
    }
  });
  // work items: chunks 0..C-1, parts C..C+P-1.  Dependencies: a part waits
  // for the chunks its chunk reads, a chunk with parts for its parts, any
  // other chunk for the chunks it reads.  Dequeue order: by depth in the
  // chunk graph, parts before their chunk, then in sweep order.  And the
  // critical path estimate.
  IntVec chunk_of(n);
  par_each((size_t)C, [&](size_t c) {
    for (int i = b[c]; i < b[c + 1]; i++) chunk_of[i] = (int)c;
  });
  std::vector<int> depth(C, 0);
  std::vector<double> finish(C, 0.0);
  std::vector<std::vector<int>> deps(C);
  par_each((size_t)C, [&](size_t c) {
    std::vector<char> seen(C, 0);
    std::vector<int> &d = deps[c];
    for (long long e = xbase[c]; e < xbase[c + 1]; e++) {
      const int q = chunk_of[out.x_ci[e]];
      if (!seen[q]) {
        seen[q] = 1;
        d.push_back(q);
      }
    }
    std::sort(d.begin(), d.end());
  });
  const int I = C + P;
  out.dep_ptr.assign(I + 1, 0);
  for (int c = 0; c < C; c++)
    out.dep_ptr[c + 1] = part0[c + 1] > part0[c] ? part0[c + 1] - part0[c] : (int)deps[c].size();
  for (int p = 0; p < P; p++) out.dep_ptr[C + p + 1] = (int)deps[out.p_chunk[p]].size();
  for (int it = 0; it < I; it++) out.dep_ptr[it + 1] += out.dep_ptr[it];
  out.dep.resize(out.dep_ptr[I]);
  for (int c = 0; c < C; c++) {
    int *d = &out.dep[out.dep_ptr[c]];
    if (part0[c + 1] > part0[c]) {
      for (int p = part0[c]; p < part0[c + 1]; p++) *d++ = C + p;
    } else {
      for (int q : deps[c]) *d++ = q;
    }
  }
  for (int p = 0; p < P; p++)
    std::copy(deps[out.p_chunk[p]].begin(), deps[out.p_chunk[p]].end(),
              out.dep.begin() + out.dep_ptr[C + p]);
  double crit = 0.0;
  for (int t = 0; t < C; t++) {
    const int c = backward ? C - 1 - t : t;
    int dep = 0;
    double start = 0.0;
    for (int q : deps[c]) {
      dep = std::max(dep, depth[q] + 1);
      start = std::max(start, finish[q]);
    }
    depth[c] = dep;
    const long long E = xbase[c + 1] - xbase[c];
    const int np = part0[c + 1] - part0[c];
    const double per = np ? (double)E / np + kChunkThreads : (double)E;  // entries of the slowest run
    const double batches = std::ceil(per / (kChunkThreads * kChunkExtBatch));
    finish[c] = start + (np ? 2 : 1) * kChunkFixedUs + batches * kChunkGatherUs +
                rounds_of[c] * kChunkRoundUs;
    crit = std::max(crit, finish[c]);
  }
  out.est_us = crit;
  out.order.resize(I);
  for (int it = 0; it < I; it++) out.order[it] = it;
  auto key_chunk = [&](int it) { return it < C ? it : out.p_chunk[it - C]; };
  std::stable_sort(out.order.begin(), out.order.end(), [&](int x, int y) {
    const int cx = key_chunk(x), cy = key_chunk(y);
    if (depth[cx] != depth[cy]) return depth[cx] < depth[cy];
    const int kx = x >= C ? 0 : 1, ky = y >= C ? 0 : 1;  // parts before chunks
    if (kx != ky) return kx < ky;
    if (cx != cy) return backward ? cx > cy : cx < cy;
    return x < y;
  });
}
