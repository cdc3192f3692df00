// Solve kernels for gfx950 (wave64).  All are HBM-bandwidth kernels: no
// dense contraction exists on this path, so none uses MFMA.
//
// Index arrays are int32 (the host checks the range), values FP64 or FP32.
// A row group is L consecutive lanes; a 64-lane wave holds 64/L row groups.
#pragma once
#include "device.hpp"

namespace rchol_amd {

// ---------------------------------------------------------------------------
// Sparse triangular sweep, pull form, persistent and level-scheduled.
//
// Rows are stored in "slot" order: sorted by dependency level (and inside a
// level by decreasing length), so a slot range of one level holds mutually
// independent rows.  The host cuts each level into tasks of at most 64
// slots.  A task either has a lane width L = 2^lg (lg in [0, 6]) from the
// length of its longest row, or -- rows of a wide level -- is a stream task
// (sweep_stream).  The host gives every wave a static list of tasks in
// level order (wave_ptr / wt_s0 / wt_info; no atomics).  A row reads x[j] of its dependencies; unpublished
// entries still hold the sentinel and are polled until the producing wave
// stores the value (agent-scope atomic store).  Deadlock freedom: every
// launched wave is resident (the grid is sized from the occupancy query),
// each wave walks its list in level order and a task only depends on rows
// of lower levels, so among the unfinished tasks of the lowest level, the
// one its wave is on has all inputs.  That all workgroups are resident is
// checked (grid_arrive / wave_resident), not assumed.
//
//   x[row] = (rhs[row] - sum_e val[e] * x[col[e]]) / diag[slot]
//
// A row group of L lanes takes its row in rounds of kGather * L entries:
// every lane loads its kGather indices and values, then issues all
// kGather gathers before looking at any of them, so a round costs one
// index latency and one gather latency whatever its length.  The lane
// order of the products (entry e0 + sub, + L, + 2L, ...) and the group sum
// do not depend on kGather.
//
// FUSE == 1: also store the sentinel into aux[row] (prepares the output of
//            the following sweep, which must start from sentinels).
// FUSE == 2: accumulate dotv[row] * x[row] and reduce it on the grid into
//            sc->rz; the last block forms beta (pcg.cpp:93-95).
// ---------------------------------------------------------------------------
#ifndef RCHOL_AMD_GATHER
#define RCHOL_AMD_GATHER 8
#endif
constexpr int kGather = RCHOL_AMD_GATHER;  // gathers in flight per lane and round
#ifndef RCHOL_AMD_PROXY
#define RCHOL_AMD_PROXY 1
#endif

// Sum of val[e] * x[ci[e]] over the entries e1 - 1 - sub - k L >= e0 of one
// row, waiting for unpublished x.  The row is read from its end: its entries
// are stored latest dependency last, so slot 0 of the first round -- the one
// await_batch polls -- holds the value produced last.  Wave-converged call
// (every lane of the wave, e0 == e1 for idle groups).
//
// px = pc | g << 9 with pc > 0: the first round holds a proxy
// (await_proxy) at distance pc from the end of the row of row group g, i.e.
// in lane g L + pc % L, slot pc / L; the task cut only sets it when
// pc < kGather * L.
template <typename VT, typename XT, int L, bool WAIT>
__device__ __forceinline__ XT row_dot(int e0, int e1, int sub, const int *__restrict__ ci,
                                      const VT *__restrict__ cv, const XT *x, int *err,
                                      int coherent, int px = 0) {
  XT acc = 0;
  for (int c = e1 - 1 - sub; c >= e0; c -= kGather * L) {
    int j[kGather];
    VT v[kGather];
    XT xv[kGather];
#pragma unroll
    for (int u = 0; u < kGather; u++) {
      const int e = c - u * L;
      j[u] = e >= e0 ? ld_stream(ci + e) : -1;
      v[u] = e >= e0 ? ld_stream(cv + e) : VT(0);
    }
#pragma unroll
    for (int u = 0; u < kGather; u++) {
      if (j[u] < 0)
        xv[u] = XT(0);
      else if (WAIT && coherent)
        xv[u] = Sent<XT>::load_agent(x + j[u]);
      else
        xv[u] = x[j[u]];
    }
    if (WAIT) {
      bool pend = false;
#pragma unroll
      for (int u = 0; u < kGather; u++) pend |= Sent<XT>::is(xv[u]);
      if (__ballot(pend)) {
        const int pc = px & 511;
        if (RCHOL_AMD_PROXY && pc > 0)
          await_proxy<XT, kGather>(x, j, xv, err, (px >> 9) * L + pc % L, pc / L);
        else
          await_batch<XT, kGather>(x, j, xv, err);
      }
      px = 0;  // later rounds: entries of lower levels
    }
#pragma unroll
    for (int u = 0; u < kGather; u++) acc += (XT)v[u] * xv[u];
  }
  return acc;
}

// The rows of one task (cnt <= 64 slots, metadata already in lane k of the
// m_* registers for slot k) with row groups of L lanes.
template <typename VT, typename XT, typename BT, int L, int FUSE>
__device__ __forceinline__ void sweep_task(int cnt, int m_row, int m_e0, int m_e1, VT m_dg,
                                           const int *__restrict__ ci, const VT *__restrict__ cv,
                                           const BT *__restrict__ rhs, XT *x,
                                           XT *__restrict__ aux, const double *__restrict__ dotv,
                                           int *err, int coherent, double &dacc, int px) {
  constexpr int R = 64 / L;
  const int lane = threadIdx.x & 63;
  const int grp = lane / L, sub = lane % L;
  for (int base = 0; base < cnt; base += R) {
    const int k = base + grp;  // < 64: base <= cnt - 1 - ((cnt - 1) % R)
    const int row = __shfl(m_row, k, 64);
    int e0 = __shfl(m_e0, k, 64), e1 = __shfl(m_e1, k, 64);
    const VT dg = __shfl(m_dg, k, 64);
    const bool live = k < cnt;
    if (!live) e1 = e0;
    const BT b = (live && sub == 0) ? ld_stream(rhs + row) : (BT)0;
    XT acc = row_dot<VT, XT, L, true>(e0, e1, sub, ci, cv, x, err, coherent, base == 0 ? px : 0);
    acc = group_sum<L>(acc);
    if (live && sub == 0) {
      const XT v = ((XT)b - acc) / (XT)dg;
      Sent<XT>::store_agent(x + row, v);
      if (FUSE == 1) aux[row] = Sent<XT>::value();
      if (FUSE == 2) dacc += dotv[row] * (double)v;
    }
  }
}

// A stream task (task_lg == kTaskStream): up to 64 short rows whose
// entries [E0, E1) are contiguous in slot order.  The wave loads the entries
// as one coalesced block (kStream per lane and round), issues every gather
// of the round before looking at any, parks the products in LDS and lets
// lane k add up row k's products in entry order.  One round covers
// 64 * kStream entries, so a task of a wide level costs one index latency
// and one gather latency instead of one of each per row pass.
#ifndef RCHOL_AMD_STREAM
#define RCHOL_AMD_STREAM 8
#endif
constexpr int kStream = RCHOL_AMD_STREAM;  // entries per lane and round of a stream task
constexpr int kTaskStream = 7;             // lane code of a stream task
constexpr int kTaskFlat = 8;               // lane code of a flat task (see sweep_flat)
constexpr int kFlatPerLane = 8;            // slots per lane of a flat task
constexpr int kTaskInfoShift = 10;         // task info word: slot count | lane code << 10
constexpr int kTaskProxyShift = 14;        //   | proxy (lane tasks: distance | row group << 9) << 14

// acc + buf[lo] + buf[lo + 1] + ... + buf[hi - 1], added in that order, with
// four LDS reads in flight at a time (a plain loop waits for each read).
template <typename T>
__device__ __forceinline__ T lds_row_sum(const T *buf, int lo, int hi, T acc) {
  int e = lo;
  for (; e + 4 <= hi; e += 4) {
    const T a0 = buf[e], a1 = buf[e + 1], a2 = buf[e + 2], a3 = buf[e + 3];
    acc += a0;
    acc += a1;
    acc += a2;
    acc += a3;
  }
  for (; e < hi; e++) acc += buf[e];
  return acc;
}

template <typename VT, typename XT, typename BT, int FUSE>
__device__ __forceinline__ void sweep_stream(int cnt, int m_row, int m_e0, int m_e1, VT m_dg,
                                             const int *__restrict__ ci,
                                             const VT *__restrict__ cv,
                                             const BT *__restrict__ rhs, XT *x,
                                             XT *__restrict__ aux,
                                             const double *__restrict__ dotv, int *err,
                                             int coherent, double &dacc, XT *buf) {
  const int lane = threadIdx.x & 63;
  const bool live = lane < cnt;
  const int E0 = __shfl(m_e0, 0, 64), E1 = __shfl(m_e1, cnt - 1, 64);
  const BT b = live ? ld_stream(rhs + m_row) : (BT)0;
  XT acc = 0;
  for (int c = E0; c < E1; c += 64 * kStream) {
    int j[kStream];
    VT v[kStream];
    XT xv[kStream];
#pragma unroll
    for (int u = 0; u < kStream; u++) {
      const int e = c + u * 64 + lane;
      j[u] = e < E1 ? ld_stream(ci + e) : -1;
      v[u] = e < E1 ? ld_stream(cv + e) : VT(0);
    }
#pragma unroll
    for (int u = 0; u < kStream; u++) {
      if (j[u] < 0)
        xv[u] = XT(0);
      else if (coherent)
        xv[u] = Sent<XT>::load_agent(x + j[u]);
      else
        xv[u] = x[j[u]];
    }
    bool pend = false;
#pragma unroll
    for (int u = 0; u < kStream; u++) pend |= Sent<XT>::is(xv[u]);
    if (__ballot(pend)) await_batch<XT, kStream>(x, j, xv, err);
#pragma unroll
    for (int u = 0; u < kStream; u++) buf[u * 64 + lane] = (XT)v[u] * xv[u];
    // LDS requests of one wave complete in order; the fences only keep the
    // compiler from moving the reads above the writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (live) {
      const int lo = max(m_e0, c) - c, hi = min(m_e1, c + 64 * kStream) - c;
      acc = lds_row_sum(buf, lo, hi, acc);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (live) {
    const XT r = ((XT)b - acc) / (XT)m_dg;
    Sent<XT>::store_agent(x + m_row, r);
    if (FUSE == 1) aux[m_row] = Sent<XT>::value();
    if (FUSE == 2) dacc += dotv[m_row] * (double)r;
  }
}

// A flat task (task_lg == kTaskFlat): cnt <= 64 * kFlatPerLane slots from
// s0 of a level whose rows read nothing (no off-diagonal entries: the rows
// of level 0).  Every slot's row, pivot and right-hand side are loaded at
// once, kFlatPerLane per lane, instead of one 64-row task after another.
// Same arithmetic as a row with no entries in the other task kinds.
template <typename VT, typename XT, typename BT, int FUSE>
__device__ __forceinline__ void sweep_flat(int s0, int cnt, const int *__restrict__ slot_row,
                                           const VT *__restrict__ diag,
                                           const BT *__restrict__ rhs, XT *x,
                                           XT *__restrict__ aux,
                                           const double *__restrict__ dotv, double &dacc) {
  const int lane = threadIdx.x & 63;
  int row[kFlatPerLane];
  VT dg[kFlatPerLane];
  BT b[kFlatPerLane];
#pragma unroll
  for (int u = 0; u < kFlatPerLane; u++) {
    const int t = u * 64 + lane;
    row[u] = t < cnt ? ld_stream(slot_row + s0 + t) : -1;
    dg[u] = t < cnt ? ld_stream(diag + s0 + t) : VT(1);
  }
#pragma unroll
  for (int u = 0; u < kFlatPerLane; u++) b[u] = row[u] >= 0 ? ld_stream(rhs + row[u]) : (BT)0;
#pragma unroll
  for (int u = 0; u < kFlatPerLane; u++) {
    if (row[u] < 0) continue;
    const XT v = ((XT)b[u] - XT(0)) / (XT)dg[u];
    Sent<XT>::store_agent(x + row[u], v);
    if (FUSE == 1) aux[row[u]] = Sent<XT>::value();
    if (FUSE == 2) dacc += dotv[row[u]] * (double)v;
  }
}

template <typename VT, typename XT, typename BT, int FUSE>
__global__ __launch_bounds__(256) void k_sptrsv(
    const int *__restrict__ wave_ptr, const int *__restrict__ wt_s0,
    const int *__restrict__ wt_info, const int *__restrict__ slot_row,
    const int *__restrict__ rp, const int *__restrict__ ci, const VT *__restrict__ cv,
    const VT *__restrict__ diag, const BT *__restrict__ rhs, XT *x, XT *__restrict__ aux,
    const double *__restrict__ dotv, double *__restrict__ partials, DevScalars *sc,
    const int *stop, int defer, int coherent) {
  if (*stop) return;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int *err = &sc->err;
  double dacc = 0.0;
  __shared__ XT sbuf[4][64 * kStream];  // one stream-task product buffer per wave
  XT *buf = sbuf[(threadIdx.x >> 6) & 3];
  const unsigned long long arrivals = grid_arrive(&sc->arrive);
  bool checked = false;
  const int k1 = wave_ptr[wave + 1];
  // Software pipeline over the wave's list: the descriptor and per-slot
  // metadata of task k+1 are loaded before task k runs, so a task starts
  // with its rows, entry ranges and pivots already in registers.
  int k = wave_ptr[wave];
  int n_info = 0, n_s0 = 0, n_row = 0, n_e0 = 0, n_e1 = 0;
  VT n_dg = 1;
  auto fetch = [&](int kk) {
    n_s0 = wt_s0[kk];
    n_info = wt_info[kk];
    if (((n_info >> kTaskInfoShift) & 15) != kTaskFlat &&
        lane < (n_info & ((1 << kTaskInfoShift) - 1))) {
      n_row = ld_stream(slot_row + n_s0 + lane);
      n_e0 = ld_stream(rp + n_s0 + lane);
      n_e1 = ld_stream(rp + n_s0 + lane + 1);
      n_dg = ld_stream(diag + n_s0 + lane);
    }
  };
  if (k < k1) fetch(k);
  for (; k < k1; k++) {
    const int info = n_info, m_s0 = n_s0;
    const int lg = (info >> kTaskInfoShift) & 15;
    const int cnt = info & ((1 << kTaskInfoShift) - 1);  // <= 64 slots (flat: 64 kFlatPerLane)
    const int px = info >> kTaskProxyShift;  // lane tasks' proxy (row_dot)
    const int m_row = n_row, m_e0 = n_e0, m_e1 = n_e1;
    const VT m_dg = n_dg;
    if (!checked && lg != kTaskFlat) {  // the first task that may wait
      checked = true;
      if (!wave_resident(&sc->arrive, arrivals, err)) break;
    }
    if (k + 1 < k1) fetch(k + 1);
#define RCHOL_SWEEP_TASK(LL)                                                                  \
  sweep_task<VT, XT, BT, LL, FUSE>(cnt, m_row, m_e0, m_e1, m_dg, ci, cv, rhs, x, aux, dotv, err, \
                                   coherent, dacc, px)
    switch (lg) {
      case kTaskFlat:  // forward sweeps only (FUSE 0 / 1): keeps the r.z kernel's registers
        if constexpr (FUSE != 2)
          sweep_flat<VT, XT, BT, FUSE>(m_s0, cnt, slot_row, diag, rhs, x, aux, dotv, dacc);
        else
          atomicOr(err, 2);  // not in a backward schedule
        break;
      case kTaskStream:
        sweep_stream<VT, XT, BT, FUSE>(cnt, m_row, m_e0, m_e1, m_dg, ci, cv, rhs, x, aux, dotv, err,
                                       coherent, dacc, buf);
        break;
      case 0: RCHOL_SWEEP_TASK(1); break;
      case 1: RCHOL_SWEEP_TASK(2); break;
      case 2: RCHOL_SWEEP_TASK(4); break;
      case 3: RCHOL_SWEEP_TASK(8); break;
      case 4: RCHOL_SWEEP_TASK(16); break;
      case 5: RCHOL_SWEEP_TASK(32); break;
      default: RCHOL_SWEEP_TASK(64); break;
    }
#undef RCHOL_SWEEP_TASK
  }
  if (FUSE == 2) {
    double v[1] = {dacc}, tot[1];
    if (grid_reduce<1>(v, partials, &sc->ctr[CTR_BWD], tot, sc->strict)) {
      if (defer) {
        sc->rz_acc += tot[0];  // a serial tail phase finalises
      } else {
        const double rz = sc->rz_acc + tot[0];
        sc->rz_acc = 0.0;
        sc->rz = rz;
        sc->beta = (sc->k > 0) ? rz / sc->rz_old : 0.0;
        sc->rz_old = rz;
      }
    }
  }
}

// Rows [s0, s1) of one level in a 1024-thread workgroup, row groups of L
// lanes; plain loads (every input is complete).
template <typename VT, typename XT, typename BT, int L, int FUSE>
__device__ __forceinline__ void serial_level(int s0, int s1, const int *__restrict__ slot_row,
                                             const int *__restrict__ rp,
                                             const int *__restrict__ ci,
                                             const VT *__restrict__ cv,
                                             const VT *__restrict__ diag,
                                             const BT *__restrict__ rhs, XT *x,
                                             XT *__restrict__ aux,
                                             const double *__restrict__ dotv, double &dacc) {
  const int G = blockDim.x / L;
  const int g = threadIdx.x / L, sub = threadIdx.x % L;
  for (int sb = s0; sb < s1; sb += G) {
    const int s = sb + g;
    const bool live = s < s1;
    int row = 0, e0 = 0, e1 = 0;
    if (live) {
      row = slot_row[s];
      e0 = rp[s];
      e1 = rp[s + 1];
    }
    XT acc = row_dot<VT, XT, L, false>(e0, e1, sub, ci, cv, x, nullptr, 0);
    acc = group_sum<L>(acc);
    if (live && sub == 0) {
      const XT v = ((XT)rhs[row] - acc) / (XT)diag[s];
      x[row] = v;
      if (FUSE == 1) aux[row] = Sent<XT>::value();
      if (FUSE == 2) dacc += dotv[row] * (double)v;
    }
  }
}

// Narrow levels in one workgroup: levels [lev0, lev1) of a sweep, one after
// the other, separated by workgroup barriers.  It runs as its own launch
// before (head) or after (tail) the persistent kernel, so every dependency
// outside the range is complete at launch and every one inside it was
// written by this workgroup before a barrier: plain loads, no polling, no
// cross-CU hop per level.  level_lg[l] is the lane width of level l.
// FUSE == 1: sentinel into aux.  FUSE == 2: r.z partial of these rows;
// fin == 0 adds it to sc->rz_acc (a later phase finalises), fin == 1 adds
// sc->rz_acc and finalises rz / beta (pcg.cpp:93-95).
template <typename VT, typename XT, typename BT, int FUSE>
__global__ __launch_bounds__(1024) void k_sptrsv_serial(
    int lev0, int lev1, const int *__restrict__ level_ptr, const int *__restrict__ level_lg,
    const int *__restrict__ slot_row, const int *__restrict__ rp, const int *__restrict__ ci,
    const VT *__restrict__ cv, const VT *__restrict__ diag, const BT *__restrict__ rhs, XT *x,
    XT *__restrict__ aux, const double *__restrict__ dotv, DevScalars *sc, const int *stop,
    int fin) {
  if (*stop) return;
  double dacc = 0.0;
  for (int l = lev0; l < lev1; l++) {
    const int s0 = level_ptr[l], s1 = level_ptr[l + 1];
#define RCHOL_SERIAL_LEVEL(LL)                                                                \
  serial_level<VT, XT, BT, LL, FUSE>(s0, s1, slot_row, rp, ci, cv, diag, rhs, x, aux, dotv, dacc)
    switch (level_lg[l]) {
      case 0: RCHOL_SERIAL_LEVEL(1); break;
      case 1: RCHOL_SERIAL_LEVEL(2); break;
      case 2: RCHOL_SERIAL_LEVEL(4); break;
      case 3: RCHOL_SERIAL_LEVEL(8); break;
      case 4: RCHOL_SERIAL_LEVEL(16); break;
      case 5: RCHOL_SERIAL_LEVEL(32); break;
      default: RCHOL_SERIAL_LEVEL(64); break;
    }
#undef RCHOL_SERIAL_LEVEL
    __syncthreads();
  }
  if (FUSE == 2) {
    double v[1] = {dacc};
    block_sum<1>(v);
    if (threadIdx.x == 0) {
      if (fin) {
        const double rz = sc->rz_acc + v[0];
        sc->rz_acc = 0.0;
        sc->rz = rz;
        sc->beta = (sc->k > 0) ? rz / sc->rz_old : 0.0;
        sc->rz_old = rz;
      } else {
        sc->rz_acc += v[0];
      }
    }
  }
}

// Fill a vector with the sentinel (before a sweep) -- a kernel, not a memset,
// so that it can sit in a captured graph.
template <typename T>
__global__ void k_fill_sentinel(int n, T *x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] = Sent<T>::value();
}

__global__ void k_widen(int n, const float *__restrict__ a, double *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}

// Row permutation of K-wide interleaved blocks: out[i] = in[perm[i]]
// (gather, the reference's reorder(x, P, xp), util.hpp:147-155) or
// out[perm[i]] = in[i] (scatter, y[p] = x of ex_laplace.py:30-31).
template <int GATHER>
__global__ void k_permute(int n, int K, const int *__restrict__ perm, const double *__restrict__ in,
                          double *__restrict__ out) {
  const size_t total = (size_t)n * K;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t / K, c = t % K;
    const size_t p = (size_t)perm[i] * K + c;
    if (GATHER)
      out[t] = in[p];
    else
      out[p] = in[t];
  }
}

// Plain streaming copy (16-byte vectors): the reference kernel for the
// achievable HBM bandwidth and for scaling the profiler's byte counters.
__global__ __launch_bounds__(256) void k_copy(size_t n16, const double2 *__restrict__ a,
                                              double2 *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// Gradient of a loss through x = A(w)^-1 b with respect to the weight w_e of
// edge (u_e, v_e) of A = sum_e w_e (e_u - e_v)(e_u - e_v)^T + D:
// dL/dw_e = -(lam_u - lam_v)(x_u - x_v), lam = A^-1 dL/dx.
__global__ void k_edge_grad(size_t m, const long long *__restrict__ u, const long long *__restrict__ v,
                            const double *__restrict__ lam, const double *__restrict__ x,
                            double *__restrict__ g) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
       e += (size_t)gridDim.x * blockDim.x) {
    const long long a = u[e], b = v[e];
    g[e] = -(lam[a] - lam[b]) * (x[a] - x[b]);
  }
}

// ---------------------------------------------------------------------------
// PCG vector stages.  Grid-stride, a fixed grid so that the last-block
// reductions see a bounded number of partials.
// ---------------------------------------------------------------------------

// r = b, x = 0 (or keep), y = sentinel, ||b||^2; sets the loop test of
// pcg.cpp:82 for k = 0.
template <typename YT>
__global__ __launch_bounds__(256) void k_init(int n, const double *__restrict__ b, double *r,
                                              double *x, int keep_x, YT *y, double *partials,
                                              DevScalars *sc) {
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double bi = b[i];
    r[i] = bi;
    if (!keep_x) x[i] = 0.0;
    y[i] = Sent<YT>::value();
    acc += bi * bi;
  }
  double v[1] = {acc}, tot[1];
  if (grid_reduce<1>(v, partials, &sc->ctr[CTR_INIT], tot, sc->strict)) {
    sc->bb = tot[0];
    sc->rr = tot[0];
    sc->k = 0;
    sc->err = 0;
    sc->rz_acc = 0.0;
    sc->beta = 0.0;
    sc->done = !(sqrt(tot[0]) > sqrt(tot[0]) * sc->tol && 0 < sc->maxit);
  }
}

// p_new = z + beta p_old (p = z when k == 0, pcg.cpp:87-96), q = A p_new
// (pcg.cpp:100), and the two dots p.r, p.q (pcg.cpp:101-103) -> alpha.
// A row's neighbours p_new[j] are recomputed from z and p_old (bitwise the
// same value), so p_new need not be complete before q is formed.
template <typename ZT, int L>
__global__ __launch_bounds__(256) void k_spmv_p(int n, const int *__restrict__ rp,
                                                const int *__restrict__ ci,
                                                const double *__restrict__ av,
                                                const ZT *__restrict__ z,
                                                const double *__restrict__ p_old,
                                                double *__restrict__ p_new, double *__restrict__ q,
                                                const double *__restrict__ r, double *partials,
                                                DevScalars *sc) {
  if (sc->done) return;
  constexpr int R = 64 / L;
  const double beta = sc->beta;
  const bool first = (sc->k == 0);
  const int lane = threadIdx.x & 63, grp = lane / L, sub = lane % L;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (blockDim.x >> 6);
  double apr = 0.0, apq = 0.0;
  for (int row = gw * R + grp; row < n; row += nw * R) {
    const int e0 = rp[row], e1 = rp[row + 1];
    double acc = 0.0;
    for (int e = e0 + sub; e < e1; e += L) {
      const int j = ci[e];
      const double pj = first ? (double)z[j] : beta * p_old[j] + (double)z[j];
      acc += av[e] * pj;
    }
    acc = group_sum<L>(acc);
    if (sub == 0) {
      const double pi = first ? (double)z[row] : beta * p_old[row] + (double)z[row];
      p_new[row] = pi;
      q[row] = acc;
      apr += pi * r[row];
      apq += pi * acc;
    }
  }
  double v[2] = {apr, apq}, tot[2];
  if (grid_reduce<2>(v, partials, &sc->ctr[CTR_SPMVP], tot, sc->strict)) {
    sc->pr = tot[0];
    sc->pq = tot[1];
    sc->alpha = tot[0] / tot[1];
  }
}

// x += alpha p, r -= alpha q (pcg.cpp:104,108), ||r||^2, y = sentinel for
// the next forward sweep; the last block advances k and evaluates the loop
// test of pcg.cpp:82.
template <typename YT>
__global__ __launch_bounds__(256) void k_update(int n, const double *__restrict__ p,
                                                const double *__restrict__ q, double *x, double *r,
                                                YT *y, double *partials, DevScalars *sc) {
  if (sc->done) return;
  const double alpha = sc->alpha;
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    y[i] = Sent<YT>::value();
    acc += ri * ri;
  }
  double v[1] = {acc}, tot[1];
  if (grid_reduce<1>(v, partials, &sc->ctr[CTR_UPD], tot, sc->strict)) {
    const int k = sc->k + 1;
    sc->k = k;
    sc->rr = tot[0];
    sc->done = !(sqrt(tot[0]) > sqrt(sc->bb) * sc->tol && k < sc->maxit);
  }
}

// True-residual replacement for the FP32-factor variant: every
// resid_period iterations r = b - A x in FP64 replaces the recursively
// updated residual, and the loop test is re-evaluated on it.
template <int L>
__global__ __launch_bounds__(256) void k_resid_replace(int n, const int *__restrict__ rp,
                                                       const int *__restrict__ ci,
                                                       const double *__restrict__ av,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ bvec,
                                                       double *__restrict__ r, double *partials,
                                                       DevScalars *sc) {
  if (sc->done || sc->resid_period <= 0 || (sc->k % sc->resid_period) != 0) return;
  constexpr int R = 64 / L;
  const int lane = threadIdx.x & 63, grp = lane / L, sub = lane % L;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (blockDim.x >> 6);
  double arr = 0.0;
  for (int row = gw * R + grp; row < n; row += nw * R) {
    double acc = 0.0;
    for (int e = rp[row] + sub; e < rp[row + 1]; e += L) acc += av[e] * x[ci[e]];
    acc = group_sum<L>(acc);
    if (sub == 0) {
      const double ri = bvec[row] - acc;
      r[row] = ri;
      arr += ri * ri;
    }
  }
  double v[1] = {arr}, tot[1];
  if (grid_reduce<1>(v, partials, &sc->ctr[CTR_GEN], tot, sc->strict)) {
    sc->rr = tot[0];
    sc->done = !(sqrt(tot[0]) > sqrt(sc->bb) * sc->tol && sc->k < sc->maxit);
  }
}

// q = A x - b and ||q||^2 -> relres (pcg.cpp:116-118).  With bvec == null it
// is a plain y = A x.
template <int L>
__global__ __launch_bounds__(256) void k_spmv_res(int n, const int *__restrict__ rp,
                                                  const int *__restrict__ ci,
                                                  const double *__restrict__ av,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ bvec,
                                                  double *__restrict__ y, double *partials,
                                                  DevScalars *sc) {
  constexpr int R = 64 / L;
  const int lane = threadIdx.x & 63, grp = lane / L, sub = lane % L;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (blockDim.x >> 6);
  double aqq = 0.0;
  for (int row = gw * R + grp; row < n; row += nw * R) {
    const int e0 = rp[row], e1 = rp[row + 1];
    double acc = 0.0;
    for (int e = e0 + sub; e < e1; e += L) acc += av[e] * x[ci[e]];
    acc = group_sum<L>(acc);
    if (sub == 0) {
      if (bvec) acc -= bvec[row];
      y[row] = acc;
      aqq += acc * acc;
    }
  }
  if (bvec) {
    double v[1] = {aqq}, tot[1];
    if (grid_reduce<1>(v, partials, &sc->ctr[CTR_RES], tot, sc->strict)) {
      sc->qq = tot[0];
      sc->relres = sqrt(tot[0]) / sqrt(sc->bb);
    }
  }
}

// ---------------------------------------------------------------------------
// Stream forms of the SpMV stages (RCHOL_AMD_OPT_SPMV_LANES = 0): a wave
// takes 64 consecutive rows at a time, loads their entries [E0, E1) as one
// coalesced block (kSpmvStream per lane and round), issues every gather of
// the round at once, parks the products in LDS and lane k adds up row k in
// entry order.  For the short rows of a stencil matrix this is one index
// and one gather latency per 64 rows instead of one per row pass.
// ---------------------------------------------------------------------------
constexpr int kSpmvStream = 8;

// The rows a wave walks: blocks run on XCD blockIdx.x % 8 (round-robin
// dispatch; grids are a multiple of 8 blocks), so XCD x takes the
// contiguous range [n x / 8, n (x + 1) / 8) and its waves step through it 64
// rows at a time.  The neighbours gathered by one L2 are then the rows near
// its own range instead of rows spread over the whole vector.
struct RowWalk {
  int hi, first, step;
};
__device__ __forceinline__ RowWalk xcd_rows(int n) {
  const int X = (gridDim.x % 8 == 0) ? 8 : 1;
  const int x = blockIdx.x % X, jb = blockIdx.x / X, nbx = gridDim.x / X;
  const int wpb = blockDim.x >> 6;
  const int lo = (int)((long long)n * x / X);
  RowWalk w;
  w.hi = (int)((long long)n * (x + 1) / X);
  w.first = lo + (jb * wpb + (threadIdx.x >> 6)) * 64;
  w.step = nbx * wpb * 64;
  return w;
}

// sum_e av[e] * g(ci[e]) over the entries of row r0 + lane (lanes whose row
// is >= n, the end of the walk, return 0).  buf: this wave's
// 64 * kSpmvStream doubles of LDS.
template <typename G>
__device__ __forceinline__ double stream_rows(int r0, int n, const int *__restrict__ rp,
                                              const int *__restrict__ ci,
                                              const double *__restrict__ av, G g, double *buf) {
  const int lane = threadIdx.x & 63;
  const int row = r0 + lane;
  const bool live = row < n;
  const int e0 = live ? ld_stream(rp + row) : 0, e1 = live ? ld_stream(rp + row + 1) : 0;
  const int last = min(64, n - r0) - 1;
  const int E0 = __shfl(e0, 0, 64), E1 = __shfl(e1, last, 64);
  double acc = 0.0;
  for (int c = E0; c < E1; c += 64 * kSpmvStream) {
    int j[kSpmvStream];
    double a[kSpmvStream], pv[kSpmvStream];
#pragma unroll
    for (int u = 0; u < kSpmvStream; u++) {
      const int e = c + u * 64 + lane;
      j[u] = e < E1 ? ld_stream(ci + e) : -1;
      a[u] = e < E1 ? ld_stream(av + e) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSpmvStream; u++) pv[u] = j[u] >= 0 ? g(j[u]) : 0.0;
#pragma unroll
    for (int u = 0; u < kSpmvStream; u++) buf[u * 64 + lane] = a[u] * pv[u];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (live) {
      const int lo = max(e0, c) - c, hi = min(e1, c + 64 * kSpmvStream) - c;
      acc = lds_row_sum(buf, lo, hi, acc);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return acc;
}

// The p update of the stream form, p' = z + beta p (p' = z at k == 0,
// pcg.cpp:86-96), as a streaming pass of its own: the SpMV after it then
// gathers one vector (p') instead of recomputing each neighbour's p' from
// two gathers (z and p), which halves the scattered reads that bound it.
template <typename ZT>
__global__ __launch_bounds__(256) void k_p_update(int n, const ZT *__restrict__ z,
                                                  const double *__restrict__ p_old,
                                                  double *__restrict__ p_new,
                                                  const DevScalars *sc) {
  if (sc->done) return;
  const double beta = sc->beta;
  const bool first = (sc->k == 0);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p_new[i] = first ? (double)z[i] : beta * p_old[i] + (double)z[i];
}

// k_spmv_p in stream form, after k_p_update: q = A p', p'.r, p'.q and alpha
// (same results up to the order of each row sum).
__global__ __launch_bounds__(256) void k_spmv_p_s(int n, const int *__restrict__ rp,
                                                  const int *__restrict__ ci,
                                                  const double *__restrict__ av,
                                                  const double *__restrict__ p,
                                                  double *__restrict__ q,
                                                  const double *__restrict__ r, double *partials,
                                                  DevScalars *sc) {
  if (sc->done) return;
  __shared__ double sbuf[4][64 * kSpmvStream];
  double *buf = sbuf[(threadIdx.x >> 6) & 3];
  const int lane = threadIdx.x & 63;
  const RowWalk rw = xcd_rows(n);
  double apr = 0.0, apq = 0.0;
  for (int r0 = rw.first; r0 < rw.hi; r0 += rw.step) {
    const int row = r0 + lane;
    double pi = 0.0, ri = 0.0;
    if (row < rw.hi) {
      pi = p[row];
      ri = r[row];
    }
    const double acc = stream_rows(r0, rw.hi, rp, ci, av, [&](int j) { return p[j]; }, buf);
    if (row < rw.hi) {
      q[row] = acc;
      apr += pi * ri;
      apq += pi * acc;
    }
  }
  double v[2] = {apr, apq}, tot[2];
  if (grid_reduce<2>(v, partials, &sc->ctr[CTR_SPMVP], tot, sc->strict)) {
    sc->pr = tot[0];
    sc->pq = tot[1];
    sc->alpha = tot[0] / tot[1];
  }
}

// k_resid_replace in stream form.
__global__ __launch_bounds__(256) void k_resid_replace_s(int n, const int *__restrict__ rp,
                                                         const int *__restrict__ ci,
                                                         const double *__restrict__ av,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ bvec,
                                                         double *__restrict__ r,
                                                         double *partials, DevScalars *sc) {
  if (sc->done || sc->resid_period <= 0 || (sc->k % sc->resid_period) != 0) return;
  __shared__ double sbuf[4][64 * kSpmvStream];
  double *buf = sbuf[(threadIdx.x >> 6) & 3];
  const int lane = threadIdx.x & 63;
  const RowWalk rw = xcd_rows(n);
  double arr = 0.0;
  for (int r0 = rw.first; r0 < rw.hi; r0 += rw.step) {
    const int row = r0 + lane;
    const double bi = row < rw.hi ? bvec[row] : 0.0;
    const double acc = stream_rows(r0, rw.hi, rp, ci, av, [&](int j) { return x[j]; }, buf);
    if (row < rw.hi) {
      const double ri = bi - acc;
      r[row] = ri;
      arr += ri * ri;
    }
  }
  double v[1] = {arr}, tot[1];
  if (grid_reduce<1>(v, partials, &sc->ctr[CTR_GEN], tot, sc->strict)) {
    sc->rr = tot[0];
    sc->done = !(sqrt(tot[0]) > sqrt(sc->bb) * sc->tol && sc->k < sc->maxit);
  }
}

// k_spmv_res in stream form.
__global__ __launch_bounds__(256) void k_spmv_res_s(int n, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ av,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ bvec,
                                                    double *__restrict__ y, double *partials,
                                                    DevScalars *sc) {
  __shared__ double sbuf[4][64 * kSpmvStream];
  double *buf = sbuf[(threadIdx.x >> 6) & 3];
  const int lane = threadIdx.x & 63;
  const RowWalk rw = xcd_rows(n);
  double aqq = 0.0;
  for (int r0 = rw.first; r0 < rw.hi; r0 += rw.step) {
    const int row = r0 + lane;
    const double bi = (bvec && row < rw.hi) ? bvec[row] : 0.0;
    double acc = stream_rows(r0, rw.hi, rp, ci, av, [&](int j) { return x[j]; }, buf);
    if (row < rw.hi) {
      acc -= bi;
      y[row] = acc;
      aqq += acc * acc;
    }
  }
  if (bvec) {
    double v[1] = {aqq}, tot[1];
    if (grid_reduce<1>(v, partials, &sc->ctr[CTR_RES], tot, sc->strict)) {
      sc->qq = tot[0];
      sc->relres = sqrt(tot[0]) / sqrt(sc->bb);
    }
  }
}

}  // namespace rchol_amd
