// Device-side building blocks shared by the solve kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rchol_amd {

// Control block in device memory.  The whole PCG state that the host would
// otherwise hold (alpha, beta, norms, counters) lives here so that an
// iteration never needs the host.
struct DevScalars {
  double bb;        // ||b||^2
  double rr;        // ||r||^2 of the recursively updated residual
  double rz;        // r.z of the current iteration
  double rz_old;    // r.z of the previous iteration (pcg.cpp:94 d2)
  double pr;        // p.r  (pcg.cpp:101)
  double pq;        // p.q  (pcg.cpp:102)
  double alpha;
  double beta;
  double tol;
  double relres;
  double qq;        // ||A x - b||^2 (final true residual)
  double rz_acc;    // r.z accumulated by the phases of a split backward sweep
  int k;            // completed iterations
  int maxit;
  int done;         // 1 once the loop test of pcg.cpp:82 fails
  int err;          // bit 0: a triangular-solve wait hit its bound
  int zero;         // constant 0: the stop flag of a standalone sweep
  int resid_period; // true-residual replacement period (0 = off)
  int strict;       // 1: grid reductions publish partials with release/acquire fences
  unsigned int ctr[8];  // last-block counters of the reductions (self-resetting)
  unsigned long long arrive;  // arrivals at the persistent sweeps' residency check
};

// ctr[CTR_CHUNK_Q] / [CTR_CHUNK_Q + 1]: chunk queue and exit count of the
// chunked sweeps (self-resetting)
enum { CTR_INIT = 0, CTR_BWD = 1, CTR_SPMVP = 2, CTR_UPD = 3, CTR_RES = 4, CTR_GEN = 5,
       CTR_CHUNK_Q = 6 };

// Sentinel bit patterns: signalling NaNs that arithmetic never produces.
constexpr unsigned long long kSent64 = 0x7FF4DEADBEEFCAFEull;
constexpr unsigned int kSent32 = 0x7FA5DEADu;

template <typename T> struct Sent;
template <> struct Sent<double> {
  static __device__ __forceinline__ double value() { return __longlong_as_double((long long)kSent64); }
  static __device__ __forceinline__ bool is(double v) {
    return (unsigned long long)__double_as_longlong(v) == kSent64;
  }
  static __device__ __forceinline__ double load_agent(const double *p) {
    unsigned long long b = __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)b);
  }
  static __device__ __forceinline__ void store_agent(double *p, double v) {
    __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};
template <> struct Sent<float> {
  static __device__ __forceinline__ float value() { return __uint_as_float(kSent32); }
  static __device__ __forceinline__ bool is(float v) { return __float_as_uint(v) == kSent32; }
  static __device__ __forceinline__ float load_agent(const float *p) {
    unsigned int b = __hip_atomic_load((const unsigned int *)p, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
    return __uint_as_float(b);
  }
  static __device__ __forceinline__ void store_agent(float *p, float v) {
    __hip_atomic_store((unsigned int *)p, __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
};

// Loads of data that a kernel reads exactly once (matrix indices and
// values, per-slot metadata, the right-hand side): non-temporal, so that
// the L2 keeps the gathered vector lines that neighbouring rows reuse
// instead of filling with the streamed matrix.
#ifndef RCHOL_AMD_NT
#define RCHOL_AMD_NT 1
#endif
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
#if RCHOL_AMD_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

// Bound on one dependency wait, in wall-clock ticks (the constant-rate
// counter runs at ~100 MHz: ~0.67 s), far beyond any legitimate wait.  A
// broken schedule therefore ends as an error flag, never as a hang; once the
// flag is up every other wait returns at its next check.
constexpr long long kWaitTicks = 1ll << 26;
// The bound every wait applies, per device (rchol_amd_set_option
// RCHOL_AMD_OPT_WAIT_TICKS lowers it to force the error path in tests).
__device__ long long g_wait_ticks = kWaitTicks;

// Back-off cap of the polling lane of await_batch, in s_sleep(1) units
// (64 clocks each): bounds the extra delay between publication and
// observation to ~kPollCap * 64 clocks.
#ifndef RCHOL_AMD_POLL_CAP
#define RCHOL_AMD_POLL_CAP 8
#endif
constexpr int kPollCap = RCHOL_AMD_POLL_CAP;


// Completes a round of U gathers per lane, xv[u] = x[j[u]] (j[u] < 0: an
// unused slot), once some lane of the wave still holds a sentinel.  Called
// by the whole wave, converged.  Pending slots are re-read coherently first
// (a plain load may have hit a line cached before publication); while some
// remain, one lane polls the first pending address of the wave with
// back-off, and once that value is there every lane re-reads all of its
// pending slots at once (one load latency for the whole round).  A wait
// longer than g_wait_ticks raises the error flag; a raised flag zeroes what is
// still pending and returns.
// Slot u of a lane refers to x[j[u] * stride + col] (stride / col: the
// interleaved columns of a multi-right-hand-side block; 1 / 0 otherwise).
template <typename T, int U>
__device__ __forceinline__ void await_batch(const T *x, const int (&j)[U], T (&xv)[U], int *err,
                                            int stride = 1, int col = 0) {
  x += col;
  bool pend = false;
#pragma unroll
  for (int u = 0; u < U; u++)
    if (j[u] >= 0 && Sent<T>::is(xv[u])) xv[u] = Sent<T>::load_agent(x + (size_t)j[u] * stride);
#pragma unroll
  for (int u = 0; u < U; u++) pend |= j[u] >= 0 && Sent<T>::is(xv[u]);
  unsigned long long mask = __ballot(pend);
  if (!mask) return;
  const int lane = __lane_id();
  const long long t0 = wall_clock64();
  int backoff = 1;
  while (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    int jp = -1;
#pragma unroll
    for (int u = U - 1; u >= 0; u--)
      if (j[u] >= 0 && Sent<T>::is(xv[u])) jp = j[u];
    jp = __shfl(jp, leader, 64);
    if (lane == leader) {
      for (int polls = 1;; polls++) {
        if (kPollCap > 0)
          for (int i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(1);
        if (!Sent<T>::is(Sent<T>::load_agent(x + (size_t)jp * stride))) break;
        if (backoff < kPollCap) backoff <<= 1;
        if ((polls & 7) == 0 &&
            (wall_clock64() - t0 > g_wait_ticks ||
             __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          atomicOr(err, 1);
          break;
        }
      }
    }
    backoff = 1;  // progress: the next wait is likely short
    pend = false;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (j[u] >= 0 && Sent<T>::is(xv[u])) xv[u] = Sent<T>::load_agent(x + (size_t)j[u] * stride);
#pragma unroll
    for (int u = 0; u < U; u++) pend |= j[u] >= 0 && Sent<T>::is(xv[u]);
    mask = __ballot(pend);
    if (mask && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (j[u] >= 0 && Sent<T>::is(xv[u])) xv[u] = T(0);
      break;
    }
  }
}

// Back-off cap of the all-lane phase of await_proxy, in s_sleep(1) units.
#ifndef RCHOL_AMD_POLL2_CAP
#define RCHOL_AMD_POLL2_CAP 1
#endif
constexpr int kPoll2Cap = RCHOL_AMD_POLL2_CAP;

// await_batch for a round whose wave knows a proxy: slot pu of lane pl
// holds an entry a level or two below the latest of the round (the host's
// choice), so once it is published the latest levels are being computed.  Phase 1: lane pl polls the proxy with back-off (one address,
// however long the wait).  Phase 2: every lane re-reads its own pending
// slots each poll, so a value is seen one load after its publication
// instead of after a poll and a second, re-reading load.  The entries
// still pending in phase 2 are few (most of a row's entries lie below the
// latest level).  Same bound, error flag and zeroing as await_batch.
template <typename T, int U>
__device__ __forceinline__ void await_proxy(const T *x, const int (&j)[U], T (&xv)[U], int *err,
                                            int pl, int pu) {
  bool pend = false;
#pragma unroll
  for (int u = 0; u < U; u++)
    if (j[u] >= 0 && Sent<T>::is(xv[u])) xv[u] = Sent<T>::load_agent(x + j[u]);
#pragma unroll
  for (int u = 0; u < U; u++) pend |= j[u] >= 0 && Sent<T>::is(xv[u]);
  unsigned long long mask = __ballot(pend);
  if (!mask) return;
  const int lane = __lane_id();
  const long long t0 = wall_clock64();
  int jp = -1, wait = 0;
#pragma unroll
  for (int u = 0; u < U; u++)
    if (u == pu) {
      jp = j[u];
      wait = Sent<T>::is(xv[u]) ? 1 : 0;
    }
  jp = __shfl(jp, pl, 64);
  wait = __shfl(wait, pl, 64);
  if (wait && lane == pl) {
    int backoff = 1;
    for (int polls = 1;; polls++) {
      for (int i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(1);
      if (!Sent<T>::is(Sent<T>::load_agent(x + jp))) break;
      if (backoff < kPollCap) backoff <<= 1;
      if ((polls & 7) == 0 &&
          (wall_clock64() - t0 > g_wait_ticks ||
           __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        atomicOr(err, 1);
        break;
      }
    }
  }
  int backoff = 1;
  for (int polls = 1;; polls++) {
    pend = false;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (j[u] >= 0 && Sent<T>::is(xv[u])) xv[u] = Sent<T>::load_agent(x + j[u]);
#pragma unroll
    for (int u = 0; u < U; u++) pend |= j[u] >= 0 && Sent<T>::is(xv[u]);
    mask = __ballot(pend);
    if (!mask) return;
    if ((polls & 7) == 0 &&
        (wall_clock64() - t0 > g_wait_ticks ||
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      if (lane == __ffsll((long long)mask) - 1) atomicOr(err, 1);
#pragma unroll
      for (int u = 0; u < U; u++)
        if (j[u] >= 0 && Sent<T>::is(xv[u])) xv[u] = T(0);
      return;
    }
    for (int i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(1);
    if (backoff < kPoll2Cap) backoff <<= 1;
  }
}

// Sum over the L consecutive lanes of a row group (L a power of two <= 64).
template <int L>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
  return v;
}
template <int L>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of NV doubles (blockDim.x a multiple of 64, <= 1024);
// the result is valid in thread 0.  Deterministic for a fixed launch shape.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[NV][16];
#pragma unroll
  for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) red[i][w] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0.0;
      for (int k = 0; k < nw; k++) s += red[i][k];
      v[i] = s;
    }
  }
  __syncthreads();
}

// Last-block reduction: every block publishes its NV partial sums and
// counts itself in; the block that arrives last sums all partials in block
// order (deterministic) and returns true with the totals in tot[] (thread 0
// only).  The counter resets itself, so the same kernel can be replayed from
// a captured graph.
//
// Publication.  strict == 1 is the textbook form: plain stores, then an
// acquire-release counter increment (an agent-scope release writes back the
// whole L2 of the XCD, once per block).  strict == 0 (default) keeps every
// byte of the handshake in agent-scope atomics instead: a partial is written
// by a value-returning exchange, the counter increment is issued only once
// every exchange has returned (an empty asm consumes the returned values),
// and the last block reads the partials with agent-scope atomic loads.  A
// returned read-modify-write has been performed at the point where every
// agent-scope atomic of every XCD is ordered, so the last block finds all
// partials there without any cache writeback.
template <int NV>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], double *partials, unsigned int *ctr,
                                            double (&tot)[NV], int strict) {
  block_sum<NV>(v);
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    unsigned int prev;
    if (strict) {
#pragma unroll
      for (int i = 0; i < NV; i++) partials[(size_t)blockIdx.x * NV + i] = v[i];
      prev = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long got = 0;
#pragma unroll
      for (int i = 0; i < NV; i++)
        got ^= __hip_atomic_exchange(
            (unsigned long long *)(partials + (size_t)blockIdx.x * NV + i),
            (unsigned long long)__double_as_longlong(v[i]), __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT);
      asm volatile("" ::"v"(got) : "memory");
      prev = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  if (strict) __atomic_thread_fence(__ATOMIC_ACQUIRE);
  double s[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) s[i] = 0.0;
  for (unsigned int b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
#pragma unroll
    for (int i = 0; i < NV; i++)
      s[i] += strict ? partials[(size_t)b * NV + i]
                     : Sent<double>::load_agent(partials + (size_t)b * NV + i);
  }
  block_sum<NV>(s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) tot[i] = s[i];
    __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return threadIdx.x == 0;
}

// Residency check of a persistent grid.  The level-scheduled sweeps can
// only finish when all of their workgroups are resident together (a wave
// may wait for a row of any other wave), so that is checked, not assumed:
//  * grid_arrive: every workgroup counts itself in at entry (one atomic, no
//    wait) and gets the count the whole grid reaches;
//  * wave_resident: every wave, before its first task that may wait for
//    another wave, waits (one polling lane, bounded) until that count is
//    reached.  Tasks that wait for nothing (level 0 of the forward sweep)
//    run first, while the rest of the grid is still being dispatched.
// A grid that does not fit -- too many workgroups, or CUs held by other
// work -- fails the bounded wait: the error flag gets bit 2 and the wave
// skips its tasks (taking part in the kernel's grid reduction all the same).
// The counter is never reset: each launch adds exactly gridDim.x arrivals,
// so the g-th launch of a grid waits for (g + 1) * gridDim.x; the host
// zeroes it whenever the grid size changes.
__device__ __forceinline__ unsigned long long grid_arrive(unsigned long long *arrive) {
  __shared__ unsigned long long target;
  if (threadIdx.x == 0) {
    const unsigned long long v =
        __hip_atomic_fetch_add(arrive, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    target = (v / gridDim.x + 1) * gridDim.x;
  }
  __syncthreads();
  return target;
}

__device__ __forceinline__ bool wave_resident(const unsigned long long *arrive,
                                              unsigned long long target, int *err) {
  int good = 1;
  if (__lane_id() == 0) {
    int backoff = 1;
    const long long t0 = wall_clock64();
    for (int polls = 1;
         __hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;
         polls++) {
      for (int i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(1);
      if (backoff < 8) backoff <<= 1;
      if ((polls & 15) == 0 &&
          (wall_clock64() - t0 > g_wait_ticks ||
           __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        atomicOr(err, 4);
        good = 0;
        break;
      }
    }
  }
  return __shfl(good, 0, 64) != 0;
}

}  // namespace rchol_amd
