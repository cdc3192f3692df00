// Multi-right-hand-side kernels: K columns stored interleaved (row-major
// n x K, element (i, c) at i*K + c).  One row group = K lanes, lane c owns
// column c, so every matrix entry (index + value) is loaded once per row and
// applied to K columns with one coalesced K-wide vector access.  Every
// column runs its own PCG recurrence (pcg.cpp:57-127) with its own scalars
// and convergence mask; a converged column is frozen.
#pragma once
#include "kernels.hpp"

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
  int strict;  // 1: reductions publish partials with release/acquire fences
  double tol;
  unsigned int ctr[8];
  unsigned long long arrive;  // arrivals at the sweeps' residency check (grid_register)
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
// Two reductions in one kernel must use disjoint partial regions.  The
// publication handshake is grid_reduce's: value-returning atomic exchanges
// and a relaxed counter (strict == 0), or plain stores and an
// acquire-release counter (strict == 1).
template <int K>
__device__ __forceinline__ bool column_grid_reduce(double v, double *partials, unsigned int *ctr,
                                                   double &tot, int strict) {
  __shared__ double red[16][K];
  __shared__ int is_last;
  v = column_sum<K>(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane < K) red[w][lane] = v;
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += red[i][threadIdx.x];
    double *slot = partials + (size_t)blockIdx.x * K + threadIdx.x;
    if (strict) {
      *slot = s;
    } else {
      const unsigned long long got = __hip_atomic_exchange(
          (unsigned long long *)slot, (unsigned long long)__double_as_longlong(s),
          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      asm volatile("" ::"v"(got) : "memory");
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int prev =
        strict ? __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT)
               : __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  if (strict) __atomic_thread_fence(__ATOMIC_ACQUIRE);
  if (threadIdx.x < K) {
    double s = 0.0;
    for (unsigned int b = 0; b < gridDim.x; b++) {
      const double *slot = partials + (size_t)b * K + threadIdx.x;
      s += strict ? *slot : Sent<double>::load_agent(slot);
    }
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

// Level-scheduled sweep on K interleaved columns (the deadlock-freedom
// argument of k_sptrsv; per-wave task lists cut for K columns).  A row
// takes K * Le lanes: lane (grp, sub, c) = lane / (K Le), lane / K % Le,
// lane % K owns column c of entries e0 + sub, + Le, ... of row grp of the
// pass, so one index/value load serves K columns and every gather is one
// contiguous K-wide segment.  A round issues kGather gathers per lane
// before looking at any; pending values are awaited wave-cooperatively.
// FUSE 1: sentinel into aux; FUSE 2: per-column dotv.x -> rz, beta.
template <typename VT, typename XT, int Le, int K>
__device__ __forceinline__ XT row_dot_mr(int e0, int e1, int sub, int c,
                                         const int *__restrict__ ci, const VT *__restrict__ cv,
                                         const XT *x, int *err) {
  XT acc = 0;
  for (int b0 = e0 + sub; b0 < e1; b0 += kGather * Le) {
    int j[kGather];
    VT v[kGather];
    XT xv[kGather];
#pragma unroll
    for (int u = 0; u < kGather; u++) {
      const int e = b0 + u * Le;
      j[u] = e < e1 ? ld_stream(ci + e) : -1;
      v[u] = e < e1 ? ld_stream(cv + e) : VT(0);
    }
#pragma unroll
    for (int u = 0; u < kGather; u++) xv[u] = j[u] < 0 ? XT(0) : x[(size_t)j[u] * K + c];
    bool pend = false;
#pragma unroll
    for (int u = 0; u < kGather; u++) pend |= Sent<XT>::is(xv[u]);
    if (__ballot(pend)) await_batch<XT, kGather>(x, j, xv, err, K, c);
#pragma unroll
    for (int u = 0; u < kGather; u++) acc += (XT)v[u] * xv[u];
  }
  return acc;
}

template <typename VT, typename XT, typename BT, int K, int Le, int FUSE>
__device__ __forceinline__ void sweep_task_mr(int cnt, int m_row, int m_e0, int m_e1, VT m_dg,
                                              const int *__restrict__ ci,
                                              const VT *__restrict__ cv,
                                              const BT *__restrict__ rhs, XT *x,
                                              XT *__restrict__ aux,
                                              const double *__restrict__ dotv, int *err,
                                              double &dacc) {
  constexpr int W = K * Le, R = 64 / W;
  const int lane = threadIdx.x & 63;
  const int c = lane % K, sub = (lane / K) % Le, grp = lane / W;
  for (int base = 0; base < cnt; base += R) {
    const int k = base + grp;  // < 64
    const int row = __shfl(m_row, k, 64);
    int e0 = __shfl(m_e0, k, 64), e1 = __shfl(m_e1, k, 64);
    const VT dg = __shfl(m_dg, k, 64);
    const bool live = k < cnt;
    if (!live) e1 = e0;
    const size_t o = (size_t)row * K + c;
    const BT b = (live && sub == 0) ? ld_stream(rhs + o) : (BT)0;
    XT acc = row_dot_mr<VT, XT, Le, K>(e0, e1, sub, c, ci, cv, x, err);
#pragma unroll
    for (int off = K; off < W; off <<= 1) acc += __shfl_xor(acc, off, 64);
    if (live && sub == 0) {
      const XT v = ((XT)b - acc) / (XT)dg;
      Sent<XT>::store_agent(x + o, v);
      if (FUSE == 1) aux[o] = Sent<XT>::value();
      if (FUSE == 2) dacc += dotv[o] * (double)v;
    }
  }
}

template <typename VT, typename XT, typename BT, int K, int FUSE>
__global__ __launch_bounds__(256) void k_sptrsv_mr(
    const int *__restrict__ wave_ptr, const int *__restrict__ wt_s0,
    const int *__restrict__ wt_info, const int *__restrict__ slot_row,
    const int *__restrict__ rp, const int *__restrict__ ci, const VT *__restrict__ cv,
    const VT *__restrict__ diag, const BT *__restrict__ rhs, XT *x, XT *__restrict__ aux,
    const double *__restrict__ dotv, double *__restrict__ partials, BatchScalars *bs,
    const int *stop) {
  if (*stop) return;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int *err = &bs->err;
  double dacc = 0.0;
  const unsigned long long arrivals = grid_arrive(&bs->arrive);
  const int k1 = wave_ptr[wave + 1];
  int k = wave_resident(&bs->arrive, arrivals, err) ? wave_ptr[wave] : k1;
  int n_info = 0, n_row = 0, n_e0 = 0, n_e1 = 0;
  VT n_dg = 1;
  auto fetch = [&](int kk) {
    const int s = wt_s0[kk];
    n_info = wt_info[kk];
    if (lane < (n_info & ((1 << kTaskInfoShift) - 1))) {
      n_row = ld_stream(slot_row + s + lane);
      n_e0 = ld_stream(rp + s + lane);
      n_e1 = ld_stream(rp + s + lane + 1);
      n_dg = ld_stream(diag + s + lane);
    }
  };
  if (k < k1) fetch(k);
  for (; k < k1; k++) {
    const int info = n_info;
    const int lg = info >> kTaskInfoShift;
    const int cnt = info & ((1 << kTaskInfoShift) - 1);
    const int m_row = n_row, m_e0 = n_e0, m_e1 = n_e1;
    const VT m_dg = n_dg;
    if (k + 1 < k1) fetch(k + 1);
#define RCHOL_MR_TASK(LE)                                                                  \
  sweep_task_mr<VT, XT, BT, K, ((LE) * K <= 64 ? (LE) : 64 / K), FUSE>(                    \
      cnt, m_row, m_e0, m_e1, m_dg, ci, cv, rhs, x, aux, dotv, err, dacc)
    switch (lg) {
      case 0: RCHOL_MR_TASK(1); break;
      case 1: RCHOL_MR_TASK(2); break;
      case 2: RCHOL_MR_TASK(4); break;
      case 3: RCHOL_MR_TASK(8); break;
      default: RCHOL_MR_TASK(16); break;
    }
#undef RCHOL_MR_TASK
  }
  if (FUSE == 2) {
    double tot = 0.0;
    if (column_grid_reduce<K>(dacc, partials, &bs->ctr[CTR_BWD], tot, bs->strict) &&
        threadIdx.x < K) {
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
  if (column_grid_reduce<K>(acc, partials, &bs->ctr[CTR_INIT], tot, bs->strict)) {
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

// Rows of a multi-column SpMV: K lanes per row (lane c = column c), 64/K
// rows per wave step, the XCD-contiguous walk of xcd_rows in steps of
// 64/K rows.  Entries go in rounds of kGather per lane: all index/value
// loads, then all K-wide gathers, then the products.
template <int K>
__device__ __forceinline__ RowWalk xcd_rows_mr(int n) {
  constexpr int R = 64 / K;
  const int X = (gridDim.x % 8 == 0) ? 8 : 1;
  const int x = blockIdx.x % X, jb = blockIdx.x / X, nbx = gridDim.x / X;
  const int wpb = blockDim.x >> 6;
  const int lo = (int)((long long)n * x / X);
  RowWalk w;
  w.hi = (int)((long long)n * (x + 1) / X);
  w.first = lo + (jb * wpb + (threadIdx.x >> 6)) * R + (threadIdx.x & 63) / K;
  w.step = nbx * wpb * R;
  return w;
}

template <int K, typename G>
__device__ __forceinline__ double row_dot_a(int e0, int e1, const int *__restrict__ ci,
                                            const double *__restrict__ av, G g) {
  double acc = 0.0;
  for (int b0 = e0; b0 < e1; b0 += kGather) {
    int j[kGather];
    double a[kGather], pv[kGather];
#pragma unroll
    for (int u = 0; u < kGather; u++) {
      const int e = b0 + u;
      j[u] = e < e1 ? ld_stream(ci + e) : -1;
      a[u] = e < e1 ? ld_stream(av + e) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kGather; u++) pv[u] = j[u] >= 0 ? g(j[u]) : 0.0;
#pragma unroll
    for (int u = 0; u < kGather; u++) acc += a[u] * pv[u];
  }
  return acc;
}

// p_new = z + beta_c p_old (p = z at k_c == 0), q = A p_new, per-column
// p.r and p.q -> alpha_c.
// Occupancy 8 (<= 64 VGPRs): the grid fills 8 workgroups per CU, and at 7
// resident the eighth would run as a second wave of workgroups.
template <typename ZT, int K>
__global__ __launch_bounds__(256, 8) void k_spmv_p_mr(int n, const int *__restrict__ rp,
                                                   const int *__restrict__ ci,
                                                   const double *__restrict__ av,
                                                   const ZT *__restrict__ z,
                                                   const double *__restrict__ p_old,
                                                   double *__restrict__ p_new,
                                                   double *__restrict__ q,
                                                   const double *__restrict__ r,
                                                   double *partials, BatchScalars *bs) {
  if (bs->all_done) return;
  const int c = (threadIdx.x & 63) % K;
  const bool first = bs->k[c] == 0;
  const double beta = bs->beta[c];
  auto pval = [&](int j) -> double {
    const size_t o = (size_t)j * K + c;
    return first ? (double)z[o] : beta * p_old[o] + (double)z[o];
  };
  const RowWalk rw = xcd_rows_mr<K>(n);
  double apr = 0.0, apq = 0.0;
  for (int row = rw.first; row < rw.hi; row += rw.step) {
    const int e0 = ld_stream(rp + row), e1 = ld_stream(rp + row + 1);
    const size_t o = (size_t)row * K + c;
    const double pi = pval(row), ri = r[o];
    const double acc = row_dot_a<K>(e0, e1, ci, av, pval);
    p_new[o] = pi;
    q[o] = acc;
    apr += pi * ri;
    apq += pi * acc;
  }
  double tot = 0.0;
  if (column_grid_reduce<K>(apr, partials, &bs->ctr[CTR_SPMVP], tot, bs->strict) &&
      threadIdx.x < K)
    bs->pr[threadIdx.x] = tot;
  // second reduction on a disjoint partial region
  if (column_grid_reduce<K>(apq, partials + (size_t)gridDim.x * K, &bs->ctr[CTR_GEN], tot,
                            bs->strict) &&
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
  if (column_grid_reduce<K>(acc, partials, &bs->ctr[CTR_UPD], tot, bs->strict)) {
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
  const int c = (threadIdx.x & 63) % K;
  const RowWalk rw = xcd_rows_mr<K>(n);
  double aqq = 0.0;
  for (int row = rw.first; row < rw.hi; row += rw.step) {
    const int e0 = ld_stream(rp + row), e1 = ld_stream(rp + row + 1);
    const size_t o = (size_t)row * K + c;
    const double bi = bvec ? bvec[o] : 0.0;
    double acc = row_dot_a<K>(e0, e1, ci, av, [&](int j) { return x[(size_t)j * K + c]; });
    acc -= bi;
    y[o] = acc;
    aqq += acc * acc;
  }
  if (bvec) {
    double tot = 0.0;
    if (column_grid_reduce<K>(aqq, partials, &bs->ctr[CTR_RES], tot, bs->strict) &&
        threadIdx.x < K)
      bs->relres[threadIdx.x] = sqrt(tot) / sqrt(bs->bb[threadIdx.x]);
  }
}

}  // namespace rchol_amd
