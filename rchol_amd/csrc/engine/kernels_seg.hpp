// Uneven batches: S independent systems stacked as one block-diagonal A and
// G (system s owns rows [off_s, off_{s+1})).  The triangular sweeps need no
// change: the union's level schedule interleaves the systems and every launch
// covers all of them.  The vector stages run segment-major: block b works
// only on rows of system seg[b], so each block emits exactly one partial per
// dot and the last block of a system (per-system counter) reduces that
// system's partials in block order -- deterministic, no cross-system
// contention.  Every system keeps its own PCG scalars and stops on its own
// test (pcg.cpp:82); a converged system is frozen.
#pragma once
#include "kernels.hpp"

namespace rchol_amd {

constexpr int kMaxSeg = 256;

struct SegScalars {
  double bb[kMaxSeg], rr[kMaxSeg], rz_old[kMaxSeg], pr[kMaxSeg];
  double alpha[kMaxSeg], beta[kMaxSeg], relres[kMaxSeg];
  int k[kMaxSeg];
  int done[kMaxSeg];
  unsigned int ctr[6][kMaxSeg];  // per-system last-block counters per kernel
  unsigned int fin[6];           // systems finalised per kernel (for all_done)
  int nseg;
  int all_done;
  int maxit;
  int err;
  int strict;  // 1: partials published with release/acquire fences
  double tol;
};

// Block layout: block b covers rows [r0[b], r1[b]) of system seg[b]; the
// blocks of system s are [sb[s], sb[s+1]).
struct SegLayout {
  const int *seg, *r0, *r1, *sb;
};

// Sum NV values over the block, publish the block's partials, and report
// whether this block is the last of its system.  Block-uniform result.  The
// handshake is grid_reduce's: value-returning atomic exchanges and a relaxed
// counter (strict == 0; seg_total then reads with atomic loads), or plain
// stores and an acquire-release counter (strict == 1).
template <int NV>
__device__ __forceinline__ bool seg_reduce(double (&v)[NV], double *partials, unsigned int *ctr,
                                           int nblocks_in_seg, int strict) {
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
    is_last = (prev == (unsigned int)nblocks_in_seg - 1);
    if (is_last) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (is_last && strict) __atomic_thread_fence(__ATOMIC_ACQUIRE);
  return is_last;
}

// Thread 0 of a system's last block: sum the system's partials in block order.
template <int NV>
__device__ __forceinline__ void seg_total(const double *partials, int b0, int b1,
                                          double (&tot)[NV], int strict) {
#pragma unroll
  for (int i = 0; i < NV; i++) tot[i] = 0.0;
  for (int b = b0; b < b1; b++) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      const double *slot = partials + (size_t)b * NV + i;
      tot[i] += strict ? *slot : Sent<double>::load_agent(slot);
    }
  }
}

// After system s was finalised by kernel `kind`: the last system to finish
// forms all_done (AND of the masks).
__device__ __forceinline__ void seg_finish(SegScalars *ss, int kind) {
  unsigned int prev =
      __hip_atomic_fetch_add(&ss->fin[kind], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  if (prev == (unsigned int)ss->nseg - 1) {
    int all = 1;
    for (int s = 0; s < ss->nseg; s++) all &= __hip_atomic_load(&ss->done[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ss->all_done = all;
    __hip_atomic_store(&ss->fin[kind], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename YT>
__global__ __launch_bounds__(256) void k_init_seg(SegLayout lay, const double *__restrict__ b,
                                                  double *r, double *x, int keep_x, YT *y,
                                                  double *partials, SegScalars *ss) {
  const int s = lay.seg[blockIdx.x];
  double acc = 0.0;
  for (int i = lay.r0[blockIdx.x] + threadIdx.x; i < lay.r1[blockIdx.x]; i += blockDim.x) {
    const double bi = b[i];
    r[i] = bi;
    if (!keep_x) x[i] = 0.0;
    y[i] = Sent<YT>::value();
    acc += bi * bi;
  }
  double v[1] = {acc};
  const int b0 = lay.sb[s], b1 = lay.sb[s + 1];
  if (seg_reduce<1>(v, partials, &ss->ctr[0][s], b1 - b0, ss->strict) && threadIdx.x == 0) {
    double t[1];
    seg_total<1>(partials, b0, b1, t, ss->strict);
    ss->bb[s] = t[0];
    ss->rr[s] = t[0];
    ss->k[s] = 0;
    ss->beta[s] = 0.0;
    ss->done[s] = !(sqrt(t[0]) > sqrt(t[0]) * ss->tol && 0 < ss->maxit);
    if (s == 0) ss->err = 0;
    seg_finish(ss, 0);
  }
}

// r.z per system -> beta (pcg.cpp:93-95).
template <typename ZT>
__global__ __launch_bounds__(256) void k_dot_seg(SegLayout lay, const double *__restrict__ r,
                                                 const ZT *__restrict__ z, double *partials,
                                                 SegScalars *ss) {
  if (ss->all_done) return;
  const int s = lay.seg[blockIdx.x];
  double acc = 0.0;
  for (int i = lay.r0[blockIdx.x] + threadIdx.x; i < lay.r1[blockIdx.x]; i += blockDim.x)
    acc += r[i] * (double)z[i];
  double v[1] = {acc};
  const int b0 = lay.sb[s], b1 = lay.sb[s + 1];
  if (seg_reduce<1>(v, partials, &ss->ctr[1][s], b1 - b0, ss->strict) && threadIdx.x == 0) {
    double t[1];
    seg_total<1>(partials, b0, b1, t, ss->strict);
    ss->beta[s] = ss->k[s] > 0 ? t[0] / ss->rz_old[s] : 0.0;
    ss->rz_old[s] = t[0];
  }
}

// p' = z + beta_s p, q = A p', p'.r, p'.q -> alpha_s.
template <typename ZT, int L>
__global__ __launch_bounds__(256) void k_spmv_p_seg(SegLayout lay, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ av,
                                                    const ZT *__restrict__ z,
                                                    const double *__restrict__ p_old,
                                                    double *__restrict__ p_new,
                                                    double *__restrict__ q,
                                                    const double *__restrict__ r,
                                                    double *partials, SegScalars *ss) {
  if (ss->all_done) return;
  constexpr int R = 64 / L;
  const int s = lay.seg[blockIdx.x];
  const bool first = ss->k[s] == 0;
  const double beta = ss->beta[s];
  const int lane = threadIdx.x & 63, grp = lane / L, sub = lane % L;
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  double apr = 0.0, apq = 0.0;
  for (int row = lay.r0[blockIdx.x] + wv * R + grp; row < lay.r1[blockIdx.x]; row += nwv * R) {
    double acc = 0.0;
    for (int e = rp[row] + sub; e < rp[row + 1]; e += L) {
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
  double v[2] = {apr, apq};
  const int b0 = lay.sb[s], b1 = lay.sb[s + 1];
  if (seg_reduce<2>(v, partials, &ss->ctr[2][s], b1 - b0, ss->strict) && threadIdx.x == 0) {
    double t[2];
    seg_total<2>(partials, b0, b1, t, ss->strict);
    ss->pr[s] = t[0];
    ss->alpha[s] = t[0] / t[1];
  }
}

// The p update of the stream form, per system (p' = z + beta_s p; p' = z at
// its k == 0, pcg.cpp:86-96), as a streaming pass of its own over the same
// blocks: the SpMV after it gathers p' only, one scattered read per entry
// instead of two.
template <typename ZT>
__global__ __launch_bounds__(256) void k_p_update_seg(SegLayout lay, const ZT *__restrict__ z,
                                                      const double *__restrict__ p_old,
                                                      double *__restrict__ p_new,
                                                      const SegScalars *ss) {
  if (ss->all_done) return;
  const int s = lay.seg[blockIdx.x];
  const bool first = ss->k[s] == 0;
  const double beta = ss->beta[s];
  for (int i = lay.r0[blockIdx.x] + threadIdx.x; i < lay.r1[blockIdx.x]; i += blockDim.x)
    p_new[i] = first ? (double)z[i] : beta * p_old[i] + (double)z[i];
}

// k_spmv_p_seg in stream form (RCHOL_AMD_OPT_SPMV_LANES = 0), after
// k_p_update_seg: the waves of a block walk its row range 64 rows at a time
// (stream_rows), gathering p'.
// Occupancy 8 (<= 64 VGPRs): the grid fills 8 workgroups per CU, and at 7
// resident the eighth would run as a second wave of workgroups.
__global__ __launch_bounds__(256, 8) void k_spmv_p_seg_s(SegLayout lay, const int *__restrict__ rp,
                                                      const int *__restrict__ ci,
                                                      const double *__restrict__ av,
                                                      const double *__restrict__ p,
                                                      double *__restrict__ q,
                                                      const double *__restrict__ r,
                                                      double *partials, SegScalars *ss) {
  if (ss->all_done) return;
  __shared__ double sbuf[4][64 * kSpmvStream];
  double *buf = sbuf[(threadIdx.x >> 6) & 3];
  const int s = lay.seg[blockIdx.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int hi = lay.r1[blockIdx.x];
  double apr = 0.0, apq = 0.0;
  for (int r0 = lay.r0[blockIdx.x] + wv * 64; r0 < hi; r0 += nwv * 64) {
    const int row = r0 + lane;
    double pi = 0.0, ri = 0.0;
    if (row < hi) {
      pi = p[row];
      ri = r[row];
    }
    const double acc = stream_rows(r0, hi, rp, ci, av, [&](int j) { return p[j]; }, buf);
    if (row < hi) {
      q[row] = acc;
      apr += pi * ri;
      apq += pi * acc;
    }
  }
  double v[2] = {apr, apq};
  const int b0 = lay.sb[s], b1 = lay.sb[s + 1];
  if (seg_reduce<2>(v, partials, &ss->ctr[2][s], b1 - b0, ss->strict) && threadIdx.x == 0) {
    double t[2];
    seg_total<2>(partials, b0, b1, t, ss->strict);
    ss->pr[s] = t[0];
    ss->alpha[s] = t[0] / t[1];
  }
}

template <typename YT>
__global__ __launch_bounds__(256) void k_update_seg(SegLayout lay, const double *__restrict__ p,
                                                    const double *__restrict__ q, double *x,
                                                    double *r, YT *y, double *partials,
                                                    SegScalars *ss) {
  if (ss->all_done) return;
  const int s = lay.seg[blockIdx.x];
  const bool live = !ss->done[s];
  const double alpha = ss->alpha[s];
  double acc = 0.0;
  for (int i = lay.r0[blockIdx.x] + threadIdx.x; i < lay.r1[blockIdx.x]; i += blockDim.x) {
    double ri = r[i];
    if (live) {
      x[i] += alpha * p[i];
      ri = ri - alpha * q[i];
      r[i] = ri;
    }
    y[i] = Sent<YT>::value();
    acc += ri * ri;
  }
  double v[1] = {acc};
  const int b0 = lay.sb[s], b1 = lay.sb[s + 1];
  if (seg_reduce<1>(v, partials, &ss->ctr[3][s], b1 - b0, ss->strict) && threadIdx.x == 0) {
    if (!ss->done[s]) {
      double t[1];
      seg_total<1>(partials, b0, b1, t, ss->strict);
      const int k = ss->k[s] + 1;
      ss->k[s] = k;
      ss->rr[s] = t[0];
      ss->done[s] = !(sqrt(t[0]) > sqrt(ss->bb[s]) * ss->tol && k < ss->maxit);
    }
    seg_finish(ss, 3);
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_spmv_res_seg(SegLayout lay, const int *__restrict__ rp,
                                                      const int *__restrict__ ci,
                                                      const double *__restrict__ av,
                                                      const double *__restrict__ x,
                                                      const double *__restrict__ bvec,
                                                      double *__restrict__ y, double *partials,
                                                      SegScalars *ss) {
  constexpr int R = 64 / L;
  const int s = lay.seg[blockIdx.x];
  const int lane = threadIdx.x & 63, grp = lane / L, sub = lane % L;
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  double aqq = 0.0;
  for (int row = lay.r0[blockIdx.x] + wv * R + grp; row < lay.r1[blockIdx.x]; row += nwv * R) {
    double acc = 0.0;
    for (int e = rp[row] + sub; e < rp[row + 1]; e += L) acc += av[e] * x[ci[e]];
    acc = group_sum<L>(acc);
    if (sub == 0) {
      acc -= bvec[row];
      y[row] = acc;
      aqq += acc * acc;
    }
  }
  double v[1] = {aqq};
  const int b0 = lay.sb[s], b1 = lay.sb[s + 1];
  if (seg_reduce<1>(v, partials, &ss->ctr[4][s], b1 - b0, ss->strict) && threadIdx.x == 0) {
    double t[1];
    seg_total<1>(partials, b0, b1, t, ss->strict);
    ss->relres[s] = sqrt(t[0]) / sqrt(ss->bb[s]);
  }
}

// k_spmv_res_seg in stream form.
__global__ __launch_bounds__(256) void k_spmv_res_seg_s(SegLayout lay, const int *__restrict__ rp,
                                                        const int *__restrict__ ci,
                                                        const double *__restrict__ av,
                                                        const double *__restrict__ x,
                                                        const double *__restrict__ bvec,
                                                        double *__restrict__ y, double *partials,
                                                        SegScalars *ss) {
  __shared__ double sbuf[4][64 * kSpmvStream];
  double *buf = sbuf[(threadIdx.x >> 6) & 3];
  const int s = lay.seg[blockIdx.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int hi = lay.r1[blockIdx.x];
  double aqq = 0.0;
  for (int r0 = lay.r0[blockIdx.x] + wv * 64; r0 < hi; r0 += nwv * 64) {
    const int row = r0 + lane;
    const double bi = row < hi ? bvec[row] : 0.0;
    double acc = stream_rows(r0, hi, rp, ci, av, [&](int j) { return x[j]; }, buf);
    if (row < hi) {
      acc -= bi;
      y[row] = acc;
      aqq += acc * acc;
    }
  }
  double v[1] = {aqq};
  const int b0 = lay.sb[s], b1 = lay.sb[s + 1];
  if (seg_reduce<1>(v, partials, &ss->ctr[4][s], b1 - b0, ss->strict) && threadIdx.x == 0) {
    double t[1];
    seg_total<1>(partials, b0, b1, t, ss->strict);
    ss->relres[s] = sqrt(t[0]) / sqrt(ss->bb[s]);
  }
}

}  // namespace rchol_amd
