// Multi-right-hand-side kernels: K columns stored interleaved (row-major
// n x K, element (i, c) at i*K + c).  One row group = K lanes, lane c owns
// column c, so every matrix entry (index + value) is loaded once per row and
// applied to K columns with one coalesced K-wide vector access.  Every
// column runs its own PCG recurrence (pcg.cpp:57-127) with its own scalars
// and convergence mask; a converged column is frozen.
#pragma once
#include "device.hpp"

namespace rchol_amd {

constexpr int kMaxCols = 32;

struct BatchScalars {
  double bb[kMaxCols], rr[kMaxCols], rz[kMaxCols], rz_old[kMaxCols];
  double pr[kMaxCols], pq[kMaxCols], alpha[kMaxCols], beta[kMaxCols];
  double relres[kMaxCols];
  int k[kMaxCols];
  int done[kMaxCols];
  int all_done;
  int ncols;  // live columns (<= K); padding columns start done
  int maxit;
  int err;
  int zero;
  double tol;
  unsigned int ctr[8];
};

// Sum lanes holding the same column inside a wave (lanes c, c+K, c+2K, ...).
template <int K>
__device__ __forceinline__ double column_sum(double v) {
#pragma unroll
  for (int off = K; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Per-column block sums, then the last block sums every column over all
// blocks in block order (deterministic).  Returns true in every thread of the
// last block (block-uniform); threads c < K hold the total of column c in tot.
// Two reductions in one kernel must use disjoint partial regions.
template <int K>
__device__ __forceinline__ bool column_grid_reduce(double v, double *partials, unsigned int *ctr,
                                                   double &tot) {
  __shared__ double red[16][K];
  __shared__ int is_last;
  v = column_sum<K>(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane < K) red[w][lane] = v;
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += red[i][threadIdx.x];
    partials[(size_t)blockIdx.x * K + threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int prev =
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  if (threadIdx.x < K) {
    double s = 0.0;
    for (unsigned int b = 0; b < gridDim.x; b++) s += partials[(size_t)b * K + threadIdx.x];
    tot = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// AND of the column masks into all_done (last block, after its column
// threads updated done[]).
template <int K>
__device__ __forceinline__ void set_all_done(BatchScalars *bs) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int c = 0; c < K; c++) all &= bs->done[c];
    bs->all_done = all;
  }
}

// Level-scheduled sweep on K interleaved columns (same schedule and
// deadlock-freedom argument as k_sptrsv); a wave holds 64/K rows.
// FUSE 1: sentinel into aux; FUSE 2: per-column dotv.x -> rz, beta.
template <typename VT, typename XT, typename BT, int K, int FUSE>
__global__ __launch_bounds__(256) void k_sptrsv_mr(
    int ntasks, const int *__restrict__ task_ptr, const int *__restrict__ slot_row,
    const int *__restrict__ rp, const int *__restrict__ ci, const VT *__restrict__ cv,
    const VT *__restrict__ diag, const BT *__restrict__ rhs, XT *x, XT *__restrict__ aux,
    const double *__restrict__ dotv, double *__restrict__ partials, BatchScalars *bs,
    const int *stop) {
  if (*stop) return;
  constexpr int R = 64 / K;
  const int lane = threadIdx.x & 63;
  const int grp = lane / K, c = lane % K;
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int *err = &bs->err;
  double dacc = 0.0;
  for (int t = wave; t < ntasks; t += nwaves) {
    const int s0 = task_ptr[t], s1 = task_ptr[t + 1];
    const int cnt = s1 - s0;  // <= 64 slots: one coalesced metadata load
    int m_row = 0, m_e0 = 0, m_e1 = 0;
    VT m_dg = 1;
    if (lane < cnt) {
      m_row = slot_row[s0 + lane];
      m_e0 = rp[s0 + lane];
      m_e1 = rp[s0 + lane + 1];
      m_dg = diag[s0 + lane];
    }
    for (int base = 0; base < cnt; base += R) {
      const int kk = base + grp;
      const int row = __shfl(m_row, kk, 64);
      const int e0 = __shfl(m_e0, kk, 64), e1 = __shfl(m_e1, kk, 64);
      const VT dg = __shfl(m_dg, kk, 64);
      if (kk >= cnt) continue;
      const size_t o = (size_t)row * K + c;
      const BT b = rhs[o];
      XT acc = 0;
#pragma unroll 4
      for (int e = e0; e < e1; e++) {
        const int j = ci[e];
        const XT xj = wait_ready<XT>(x + (size_t)j * K + c, err);
        acc += (XT)cv[e] * xj;
      }
      const XT v = ((XT)b - acc) / (XT)dg;
      Sent<XT>::store_agent(x + o, v);
      if (FUSE == 1) aux[o] = Sent<XT>::value();
      if (FUSE == 2) dacc += dotv[o] * (double)v;
    }
  }
  if (FUSE == 2) {
    double tot = 0.0;
    if (column_grid_reduce<K>(dacc, partials, &bs->ctr[CTR_BWD], tot) && threadIdx.x < K) {
      const int cc = threadIdx.x;
      bs->rz[cc] = tot;
      bs->beta[cc] = (bs->k[cc] > 0) ? tot / bs->rz_old[cc] : 0.0;
      bs->rz_old[cc] = tot;
    }
  }
}

// r = b, x = 0 (or keep), y = sentinel, per-column ||b||^2 and loop test.
template <typename YT, int K>
__global__ __launch_bounds__(256) void k_init_mr(int n, const double *__restrict__ b, double *r,
                                                 double *x, int keep_x, YT *y, double *partials,
                                                 BatchScalars *bs) {
  double acc = 0.0;
  const size_t total = (size_t)n * K;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const double bi = b[i];
    r[i] = bi;
    if (!keep_x) x[i] = 0.0;
    y[i] = Sent<YT>::value();
    acc += bi * bi;
  }
  double tot = 0.0;
  if (column_grid_reduce<K>(acc, partials, &bs->ctr[CTR_INIT], tot)) {
    if (threadIdx.x < K) {
      const int cc = threadIdx.x;
      bs->bb[cc] = tot;
      bs->rr[cc] = tot;
      bs->k[cc] = 0;
      bs->beta[cc] = 0.0;
      bs->alpha[cc] = 0.0;
      bs->done[cc] = (cc >= bs->ncols) || !(sqrt(tot) > sqrt(tot) * bs->tol && 0 < bs->maxit);
    }
    if (threadIdx.x == 0) bs->err = 0;
    set_all_done<K>(bs);
  }
}

// p_new = z + beta_c p_old (p = z at k_c == 0), q = A p_new, per-column
// p.r and p.q -> alpha_c.
template <typename ZT, int K>
__global__ __launch_bounds__(256) void k_spmv_p_mr(int n, const int *__restrict__ rp,
                                                   const int *__restrict__ ci,
                                                   const double *__restrict__ av,
                                                   const ZT *__restrict__ z,
                                                   const double *__restrict__ p_old,
                                                   double *__restrict__ p_new,
                                                   double *__restrict__ q,
                                                   const double *__restrict__ r,
                                                   double *partials, BatchScalars *bs) {
  if (bs->all_done) return;
  constexpr int R = 64 / K;
  const int lane = threadIdx.x & 63, grp = lane / K, c = lane % K;
  const bool first = bs->k[c] == 0;
  const double beta = bs->beta[c];
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (blockDim.x >> 6);
  double apr = 0.0, apq = 0.0;
  for (int row = gw * R + grp; row < n; row += nw * R) {
    const int e0 = rp[row], e1 = rp[row + 1];
    double acc = 0.0;
    for (int e = e0; e < e1; e++) {
      const size_t o = (size_t)ci[e] * K + c;
      const double pj = first ? (double)z[o] : beta * p_old[o] + (double)z[o];
      acc += av[e] * pj;
    }
    const size_t o = (size_t)row * K + c;
    const double pi = first ? (double)z[o] : beta * p_old[o] + (double)z[o];
    p_new[o] = pi;
    q[o] = acc;
    apr += pi * r[o];
    apq += pi * acc;
  }
  double tot = 0.0;
  if (column_grid_reduce<K>(apr, partials, &bs->ctr[CTR_SPMVP], tot) && threadIdx.x < K)
    bs->pr[threadIdx.x] = tot;
  // second reduction on a disjoint partial region
  if (column_grid_reduce<K>(apq, partials + (size_t)gridDim.x * K, &bs->ctr[CTR_GEN], tot) &&
      threadIdx.x < K) {
    const int cc = threadIdx.x;
    bs->pq[cc] = tot;
    bs->alpha[cc] = bs->pr[cc] / tot;
  }
}

// x += alpha_c p, r -= alpha_c q for live columns; ||r_c||^2; y = sentinel;
// the last block advances k_c and re-evaluates every column's loop test.
template <typename YT, int K>
__global__ __launch_bounds__(256) void k_update_mr(int n, const double *__restrict__ p,
                                                   const double *__restrict__ q, double *x,
                                                   double *r, YT *y, double *partials,
                                                   BatchScalars *bs) {
  if (bs->all_done) return;
  const int c = threadIdx.x % K;
  const bool live = !bs->done[c];
  const double alpha = bs->alpha[c];
  double acc = 0.0;
  const size_t total = (size_t)n * K;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    double ri = r[i];
    if (live) {
      x[i] += alpha * p[i];
      ri = ri - alpha * q[i];
      r[i] = ri;
    }
    y[i] = Sent<YT>::value();
    acc += ri * ri;
  }
  double tot = 0.0;
  if (column_grid_reduce<K>(acc, partials, &bs->ctr[CTR_UPD], tot)) {
    if (threadIdx.x < K) {
      const int cc = threadIdx.x;
      if (!bs->done[cc]) {
        const int k = bs->k[cc] + 1;
        bs->k[cc] = k;
        bs->rr[cc] = tot;
        bs->done[cc] = !(sqrt(tot) > sqrt(bs->bb[cc]) * bs->tol && k < bs->maxit);
      }
    }
    set_all_done<K>(bs);
  }
}

// Per-column true residual ||A x - b|| / ||b|| (pcg.cpp:116-118); with
// bvec == null a plain Y = A X.
template <int K>
__global__ __launch_bounds__(256) void k_spmv_res_mr(int n, const int *__restrict__ rp,
                                                     const int *__restrict__ ci,
                                                     const double *__restrict__ av,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ bvec,
                                                     double *__restrict__ y, double *partials,
                                                     BatchScalars *bs) {
  constexpr int R = 64 / K;
  const int lane = threadIdx.x & 63, grp = lane / K, c = lane % K;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (blockDim.x >> 6);
  double aqq = 0.0;
  for (int row = gw * R + grp; row < n; row += nw * R) {
    double acc = 0.0;
    for (int e = rp[row]; e < rp[row + 1]; e++) acc += av[e] * x[(size_t)ci[e] * K + c];
    const size_t o = (size_t)row * K + c;
    if (bvec) acc -= bvec[o];
    y[o] = acc;
    aqq += acc * acc;
  }
  if (bvec) {
    double tot = 0.0;
    if (column_grid_reduce<K>(aqq, partials, &bs->ctr[CTR_RES], tot) && threadIdx.x < K)
      bs->relres[threadIdx.x] = sqrt(tot) / sqrt(bs->bb[threadIdx.x]);
  }
}

}  // namespace rchol_amd
