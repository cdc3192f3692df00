// Chunked triangular sweep for factors with deep, narrow dependency chains
// (the reference's own orderings: natural order inside nested-dissection
// blocks, c++/rchol/find_separator.cpp:111-121, give tens of thousands of
// levels of a handful of rows).  The level-scheduled kernel pays one
// cross-CU handshake per level there; this one keeps whole runs of rows in
// one workgroup's LDS and pays one handshake per chunk.
//
// A chunk is a contiguous range of rows in index order (at most S rows, S
// set by the LDS size; the host cuts ranges at nested-dissection block
// boundaries found from the sparsity pattern).  Every dependency of a chunk
// row is either inside the chunk ("internal") or in a chunk that precedes it
// in sweep order ("external").  The host orders each chunk's rows by their
// dependency level inside the chunk (local index l), and a workgroup runs a
// chunk in three phases:
//   0. one thread takes the ticket and waits for the done flags of the work
//      items this one reads (a host-made list), so that waiting costs one
//      poller per workgroup, not a poll per value and thread;
//   1. all threads: yl[l] = (rhs[row] - sum_ext val * x[col]) / diag, the
//      external values read from the output vector with agent-scope loads,
//      kChunkExtBatch in flight per thread (a value still holding the
//      sentinel is polled for, as a safety net);
//   2. one wave: the internal entries, pre-divided by their row's diagonal,
//      in rounds of 64: lane k of a round applies yl[t] -= v * yl[s] with an
//      LDS atomic.  The host packs a round only with entries whose source
//      row s is final (all of its own entries sit in earlier rounds), so a
//      round's reads never see its own adds; LDS requests of one wave are
//      performed in order, so no wait separates the rounds.  Rounds are
//      prefetched kChunkPrefetch ahead into registers;
//   3. all threads: x[row] = yl[l] (agent-scope stores, each thread fenced)
//      and the fused epilogues; then the chunk's done flag (release).
// A chunk with many external entries hands their sums to "parts": work
// items of their own, run by other workgroups in parallel, each writing the
// sums of a row-aligned range of the chunk's rows into extsum; the chunk
// waits for its parts instead of the chunks they read, and starts from those
// sums.  Work items come from a queue in an order where every item follows
// the items it reads, so a workgroup only ever waits for chunks already taken by
// running workgroups: no co-residency assumption, any grid size completes.
// The queue counter and the done flags are reset by the last workgroup to
// leave, so the kernel replays from a captured graph.
//
// FUSE as in k_sptrsv: 1 stores the sentinel into aux[row], 2 reduces
// dotv . x into sc->rz and forms beta (pcg.cpp:93-95).
#pragma once
#include "device.hpp"

namespace rchol_amd {

constexpr int kChunkThreads = 512;   // threads per chunk workgroup
constexpr int kChunkPrefetch = 32;   // rounds of internal entries in flight (wave 0)
constexpr int kChunkExtBatch = 16;   // external gathers in flight per thread
constexpr unsigned kChunkPad = 0xFFFFFFFFu;  // an unused lane of a round

// x[j] once it is published: a plain load first, then agent-scope polls with
// a wall-clock bound that raises the error flag (and returns 0) instead of
// hanging.
template <typename XT>
__device__ __forceinline__ XT chunk_wait_value(const XT *x, int j, int *err) {
  XT v = Sent<XT>::load_agent(x + j);
  if (!Sent<XT>::is(v)) return v;
  const long long t0 = wall_clock64();
  for (int polls = 1;; polls++) {
    __builtin_amdgcn_s_sleep(2);
    v = Sent<XT>::load_agent(x + j);
    if (!Sent<XT>::is(v)) return v;
    if ((polls & 15) == 0 &&
        (wall_clock64() - t0 > g_wait_ticks ||
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      atomicOr(err, 1);
      return XT(0);
    }
  }
}

// Phase 2: rounds [r0, r1) of 64 packed internal entries, by one wave.
// Word of a lane: source local row (bits 0-15) | target local row << 16.
// The host pads every chunk to a whole number of kChunkPrefetch rounds and
// the arrays by kChunkPrefetch rounds past the last chunk, so the ring's
// loads need no bounds test (a test would make the compiler copy the ring
// and wait for every load).
template <typename VT, typename XT>
__device__ __forceinline__ void chunk_rounds(XT *yl, const unsigned *__restrict__ rw,
                                             const VT *__restrict__ rv, int r0, int r1) {
  const int lane = threadIdx.x & 63;
  unsigned w[kChunkPrefetch];
  VT v[kChunkPrefetch];
#pragma unroll
  for (int j = 0; j < kChunkPrefetch; j++) {
    w[j] = ld_stream(rw + (size_t)(r0 + j) * 64 + lane);
    v[j] = ld_stream(rv + (size_t)(r0 + j) * 64 + lane);
  }
  for (int r = r0; r < r1; r += kChunkPrefetch) {
#pragma unroll
    for (int j = 0; j < kChunkPrefetch; j++) {
      const unsigned cw = w[j];
      const VT cv = v[j];
      const size_t nx = (size_t)(r + j + kChunkPrefetch) * 64 + lane;
      w[j] = ld_stream(rw + nx);
      v[j] = ld_stream(rv + nx);
      if (cw != kChunkPad) {
        const XT y = yl[cw & 0xFFFFu];
        __hip_atomic_fetch_add(yl + (cw >> 16), -((XT)cv * y), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
}

// One thread's row-aligned run [e0, e1) of external entries: the sum of
// val * x[col] of each row, in entry order, handed to sink(local row, sum).
// kChunkExtBatch gathers in flight; agent-scope loads (a plain load can hit
// a line this XCD's L2 kept from before the producer published it).
template <typename VT, typename XT, typename Sink>
__device__ __forceinline__ void ext_run(int e0, int e1, const int *__restrict__ x_l,
                                        const int *__restrict__ x_ci,
                                        const VT *__restrict__ x_cv, const XT *x, int *err,
                                        Sink &&sink) {
  int cur = -1;
  XT acc = 0;
  for (int e = e0; e < e1; e += kChunkExtBatch) {
    int lr[kChunkExtBatch], j[kChunkExtBatch];
    VT v[kChunkExtBatch];
    XT xv[kChunkExtBatch];
#pragma unroll
    for (int u = 0; u < kChunkExtBatch; u++) {
      const bool ok = e + u < e1;
      lr[u] = ok ? ld_stream(x_l + e + u) : -1;
      j[u] = ok ? ld_stream(x_ci + e + u) : 0;
      v[u] = ok ? ld_stream(x_cv + e + u) : VT(0);
    }
#pragma unroll
    for (int u = 0; u < kChunkExtBatch; u++)
      xv[u] = lr[u] >= 0 ? Sent<XT>::load_agent(x + j[u]) : XT(0);
#pragma unroll
    for (int u = 0; u < kChunkExtBatch; u++)
      if (lr[u] >= 0 && Sent<XT>::is(xv[u])) xv[u] = chunk_wait_value(x, j[u], err);
#pragma unroll
    for (int u = 0; u < kChunkExtBatch; u++) {
      if (lr[u] < 0) continue;
      if (lr[u] != cur) {
        if (cur >= 0) sink(cur, acc);
        cur = lr[u];
        acc = 0;
      }
      acc += (XT)v[u] * xv[u];
    }
  }
  if (cur >= 0) sink(cur, acc);
}

// The external sums of local rows [l0, l1) of a chunk into extsum[q0 + l]
// (zero for a row without entries): row groups of L lanes, kChunkExtBatch
// gathers per lane and round, a fixed-order group sum.  Called by every
// thread of the workgroup.
template <typename VT, typename XT, int L>
__device__ __forceinline__ void ext_rows(int l0, int l1, int q0, const int *__restrict__ x_rp,
                                         const int *__restrict__ x_ci,
                                         const VT *__restrict__ x_cv, const XT *x, int *err,
                                         XT *extsum) {
  constexpr int R = 64 / L;  // rows per wave pass
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / L, sub = lane % L;
  for (int base = l0 + wave * R; base < l1; base += (kChunkThreads / 64) * R) {
    const int l = base + grp;
    const bool live = l < l1;
    const int e0 = live ? x_rp[q0 + l] : 0, e1 = live ? x_rp[q0 + l + 1] : 0;
    XT acc = 0;
    for (int c = e0 + sub; c < e1; c += L * kChunkExtBatch) {
      int j[kChunkExtBatch];
      VT v[kChunkExtBatch];
      XT xv[kChunkExtBatch];
#pragma unroll
      for (int u = 0; u < kChunkExtBatch; u++) {
        const int e = c + u * L;
        j[u] = e < e1 ? ld_stream(x_ci + e) : -1;
        v[u] = e < e1 ? ld_stream(x_cv + e) : VT(0);
      }
#pragma unroll
      for (int u = 0; u < kChunkExtBatch; u++)
        xv[u] = j[u] >= 0 ? Sent<XT>::load_agent(x + j[u]) : XT(0);
#pragma unroll
      for (int u = 0; u < kChunkExtBatch; u++)
        if (j[u] >= 0 && Sent<XT>::is(xv[u])) xv[u] = chunk_wait_value(x, j[u], err);
#pragma unroll
      for (int u = 0; u < kChunkExtBatch; u++) acc += (XT)v[u] * xv[u];
    }
    acc = group_sum<L>(acc);
    if (live && sub == 0) Sent<XT>::store_agent(extsum + q0 + l, acc);
  }
}

template <typename VT, typename XT, typename BT, int FUSE>
__global__ __launch_bounds__(kChunkThreads) void k_sptrsv_chunk(
    int nchunks, int nitems, const int *__restrict__ order, const int *__restrict__ c_q0,
    const int *__restrict__ c_m, const int *__restrict__ c_r0, const int *__restrict__ c_r1,
    const int *__restrict__ split, const int *__restrict__ c_part0,
    const int *__restrict__ p_chunk, const int *__restrict__ p_l0, const int *__restrict__ p_l1,
    const int *__restrict__ p_lg, const int *__restrict__ x_rp, const int *__restrict__ dep_ptr,
    const int *__restrict__ dep, int *done, XT *extsum, const int *__restrict__ q_row,
    const VT *__restrict__ q_diag,
    const int *__restrict__ x_l, const int *__restrict__ x_ci, const VT *__restrict__ x_cv,
    const unsigned *__restrict__ r_w, const VT *__restrict__ r_v, const BT *__restrict__ rhs,
    XT *x, XT *__restrict__ aux, const double *__restrict__ dotv, double *__restrict__ partials,
    DevScalars *sc, const int *stop) {
  if (*stop) return;
  // 16-byte aligned: the 8-byte LDS atomics of the FP64 chain need natural
  // alignment (an unaligned base, after the 4-byte s_chunk, faults)
  extern __shared__ __attribute__((aligned(16))) unsigned char chunk_smem[];
  XT *yl = (XT *)chunk_smem;
  __shared__ int s_chunk, s_last;
  int *err = &sc->err;
  unsigned *qctr = &sc->ctr[CTR_CHUNK_Q];
  const int tid = threadIdx.x;
  double dacc = 0.0;
  // Thread 0's queue work -- take a ticket, wait until the item's inputs are
  // published -- sits in one block before a barrier, at the start and at
  // the end of every item, never across the loop's back edge: with
  // thread-0 blocks on both sides of the back edge the compiler let the
  // other lanes of wave 0 loop back to the next barrier ahead of lane 0's
  // dequeue (waves then met at barriers out of step and re-ran an item).
  auto take_item = [&]() {
    const int kk =
        (int)__hip_atomic_fetch_add(qctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_chunk = kk;
    if (kk >= nitems) return;
    const int nx = order[kk];
    const long long t0 = wall_clock64();
    for (int d = dep_ptr[nx]; d < dep_ptr[nx + 1]; d++) {
      int *flag = done + dep[d];
      int backoff = 1;
      for (int polls = 1;
           !__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT); polls++) {
        for (int i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(1);
        if (backoff < 16) backoff <<= 1;
        if ((polls & 15) == 0 &&
            (wall_clock64() - t0 > g_wait_ticks ||
             __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          atomicOr(err, 1);
          break;
        }
      }
    }
  };
  if (tid == 0) take_item();
  __syncthreads();
  for (;;) {
    // wave-uniform in registers the compiler can see as uniform: the loop
    // exit and the part/chunk branch are scalar branches
    const int k = __builtin_amdgcn_readfirstlane(s_chunk);
    if (k >= nitems) break;
    const int it = __builtin_amdgcn_readfirstlane(order[k]);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (it >= nchunks) {
      // an external part: the sums of rows [l0, l1) of its chunk into extsum
      const int p = it - nchunks, c = p_chunk[p], q0 = c_q0[c];
      const int l0 = p_l0[p], l1 = p_l1[p];
      switch (__builtin_amdgcn_readfirstlane(p_lg[p])) {
        case 0: ext_rows<VT, XT, 1>(l0, l1, q0, x_rp, x_ci, x_cv, x, err, extsum); break;
        case 1: ext_rows<VT, XT, 2>(l0, l1, q0, x_rp, x_ci, x_cv, x, err, extsum); break;
        case 2: ext_rows<VT, XT, 4>(l0, l1, q0, x_rp, x_ci, x_cv, x, err, extsum); break;
        case 3: ext_rows<VT, XT, 8>(l0, l1, q0, x_rp, x_ci, x_cv, x, err, extsum); break;
        case 4: ext_rows<VT, XT, 16>(l0, l1, q0, x_rp, x_ci, x_cv, x, err, extsum); break;
        case 5: ext_rows<VT, XT, 32>(l0, l1, q0, x_rp, x_ci, x_cv, x, err, extsum); break;
        default: ext_rows<VT, XT, 64>(l0, l1, q0, x_rp, x_ci, x_cv, x, err, extsum); break;
      }
    } else {
      const int c = it;
      const int q0 = c_q0[c], m = c_m[c];
      const bool parted = c_part0[c + 1] > c_part0[c];
      // phase 1a: right-hand sides (the chunk's rows are one index range),
      // less the external sums of the chunk's parts
      for (int l = tid; l < m; l += kChunkThreads) {
        const int row = ld_stream(q_row + q0 + l);
        XT b = (XT)ld_stream(rhs + row);
        if (parted) b -= Sent<XT>::load_agent(extsum + q0 + l);
        yl[l] = b;
        if (FUSE == 1) aux[row] = Sent<XT>::value();
      }
      __syncthreads();
      // phase 1b: external entries, a row-aligned run per thread, in order
      // (empty for a chunk with parts)
      ext_run<VT, XT>(split[(size_t)c * (kChunkThreads + 1) + tid],
                      split[(size_t)c * (kChunkThreads + 1) + tid + 1], x_l, x_ci, x_cv, x, err,
                      [&](int l, XT a) { yl[l] -= a; });
      __syncthreads();
      // phase 1c: scale by the pivot (the internal entries are pre-divided)
      for (int l = tid; l < m; l += kChunkThreads)
        yl[l] = yl[l] / (XT)ld_stream(q_diag + q0 + l);
      __syncthreads();
      // phase 2: the chain, one wave
      if (tid < 64) chunk_rounds<VT, XT>(yl, r_w, r_v, c_r0[c], c_r1[c]);
      __syncthreads();
      // phase 3: publish
      for (int l = tid; l < m; l += kChunkThreads) {
        const int row = ld_stream(q_row + q0 + l);
        const XT val = yl[l];
        Sent<XT>::store_agent(x + row, val);
        if (FUSE == 2) dacc += dotv[row] * (double)val;
      }
    }
    // the item is published once every thread's stores are (both kinds end
    // here: no path of the body skips a barrier another path takes); then
    // the next ticket
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_store(done + it, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      take_item();
    }
    __syncthreads();
  }
  // the last workgroup out resets the queue and the done flags for the next
  // launch (every other workgroup has taken its last ticket by then)
  if (tid == 0)
    s_last = (int)__hip_atomic_fetch_add(qctr + 1, 1u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) == (int)gridDim.x - 1;
  __syncthreads();
  if (__builtin_amdgcn_readfirstlane(s_last)) {
    for (int k = tid; k < nitems; k += kChunkThreads)
      __hip_atomic_store(done + k, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid == 0) {
      __hip_atomic_store(qctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(qctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (FUSE == 2) {
    double v[1] = {dacc}, tot[1];
    if (grid_reduce<1>(v, partials, &sc->ctr[CTR_BWD], tot, sc->strict)) {
      const double rz = sc->rz_acc + tot[0];
      sc->rz_acc = 0.0;
      sc->rz = rz;
      sc->beta = (sc->k > 0) ? rz / sc->rz_old : 0.0;
      sc->rz_old = rz;
    }
  }
}

}  // namespace rchol_amd
