/*
 * rchol_amd -- MI355X (gfx950) solve engine for rchol's preconditioned
 * conjugate gradient.  C ABI: plain pointers and sizes, integer status codes
 * (0 = success, negative = error; rchol_amd_last_error() has the text), no
 * exceptions across the boundary.
 *
 * Matrices arrive in the reference's SparseCSR arrays (c++/sparse.hpp:26-30):
 * size_t row pointers, size_t column indices, double values.  The factor is
 * passed exactly as rchol() stores it in its SparseCSR G: the CSC arrays of
 * the lower-triangular G (rchol_lap.cpp:358-387), i.e. CSR of G^T.  The
 * preconditioner is M^-1 = G^-T G^-1 (c++/util/pcg.cpp:141-159).
 *
 * Threading: one solve per context at a time; distinct contexts may be used
 * from different host threads.
 */
#ifndef RCHOL_AMD_H
#define RCHOL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rchol_amd_ctx rchol_amd_ctx;

/* Status codes. */
#define RCHOL_AMD_OK 0
#define RCHOL_AMD_EINVAL (-1)   /* bad argument / inconsistent sizes */
#define RCHOL_AMD_EHIP (-2)     /* HIP runtime error (no GPU, OOM, launch) */
#define RCHOL_AMD_ESTATE (-3)   /* matrix or factor not uploaded */
#define RCHOL_AMD_ERANGE (-4)   /* size beyond the int32 device index range */
#define RCHOL_AMD_EPIVOT (-5)   /* zero or missing diagonal in G */
#define RCHOL_AMD_ESTALL (-6)   /* a triangular-solve wait exceeded its bound */

/* Factor storage precision for rchol_amd_upload_G. */
#define RCHOL_AMD_FP64 0
#define RCHOL_AMD_FP32 1 /* G values and the two triangular sweeps in FP32 */

/* Option keys for rchol_amd_set_option. */
#define RCHOL_AMD_OPT_GRAPH_ITERS 1   /* PCG iterations per captured hipGraph (0 = no graph) */
#define RCHOL_AMD_OPT_RESID_REPLACE 2 /* true-residual replacement period (0 = never) */
#define RCHOL_AMD_OPT_SPMV_LANES 3    /* lanes per row in SpMV (4, 8, 16 or 32), or 0
                                         (default): stream form, 64 rows per wave step */
#define RCHOL_AMD_OPT_TRSV_LANES 4    /* minimum lanes per row in the triangular sweeps
                                         (power of two <= 64, default 1); a task
                                         widens to cover its longest row in one
                                         round of 8 gathers per lane */
#define RCHOL_AMD_OPT_PROFILE 5       /* 1: time every kernel of the loop with HIP events */
#define RCHOL_AMD_OPT_SERIAL_WIDTH 6  /* leading/trailing sweep levels of at most this many
                                         rows run in one workgroup (0 = off) */
#define RCHOL_AMD_OPT_COHERENT_GATHER 7 /* 1: sweeps gather x with agent-scope loads */
#define RCHOL_AMD_OPT_TASK_WAVES 8    /* waves the sweep task cut plans for (0 = the
                                         resident capacity): fewer makes wide-level
                                         tasks (and stream tasks) appear earlier */
#define RCHOL_AMD_OPT_STRICT_SYNC 9   /* 1: the PCG loop's grid reductions publish their
                                         partials with release/acquire fences instead of
                                         value-returning atomics (same results) */
#define RCHOL_AMD_OPT_CHUNKED 10      /* single-right-hand-side sweeps in chunks of rows held
                                         in one workgroup (deep factors: the reference's own
                                         orders) instead of level by level.  -1 (default):
                                         planned at upload_G when the factor has many narrow
                                         levels, used when its estimate beats the level
                                         schedule; 0: never; 1: always (planned at the next
                                         upload_G; set before it) */
#define RCHOL_AMD_OPT_WAIT_TICKS 11   /* bound of every dependency / residency wait, in
                                         ticks of the 100 MHz wall clock, for the GPU of
                                         this context (every context on it); 0 restores
                                         the default (2^26, ~0.67 s).  A lower bound turns
                                         longer waits into RCHOL_AMD_ESTALL: for testing
                                         the error path */

/* Counters filled by rchol_amd_get_stats. */
typedef struct rchol_amd_stats {
  size_t n;
  size_t nnz_a;
  size_t nnz_g;            /* stored entries of G including the diagonal */
  int fwd_levels;          /* dependency levels of G y = r   */
  int bwd_levels;          /* dependency levels of G^T z = y */
  int fwd_tasks;
  int bwd_tasks;
  int trsv_waves;          /* resident waves of the persistent sweeps */
  int g_precision;         /* RCHOL_AMD_FP64 / RCHOL_AMD_FP32 */
  int last_iters;
  double last_solve_ms;    /* device time of the last PCG loop (events) */
  double last_relres;
  /* With RCHOL_AMD_OPT_PROFILE: summed device time of each loop kernel over
   * the last solve -- [0] forward sweep, [1] backward sweep, [2] p update +
   * SpMV, [3] x/r update -- and the iterations they cover. */
  double kernel_ms[4];
  int kernel_iters;
  /* Host wall time of the last uploads: [0] rchol_amd_upload_A (checks,
   * index narrowing, copies), and of rchol_amd_upload_G [1] the factor
   * analysis (validation, levels, slot-ordered CSR of both sweeps), [2] the
   * sweep schedule (task cut, per-wave lists), [3] the FP32 conversion (FP32
   * factor) and the copies to the device. */
  double setup_ms[4];
  /* 1 when the single-right-hand-side sweeps run chunked; chunks and rounds
   * of internal entries of [0] the forward and [1] the backward plan (0: no
   * plan). */
  int chunked;
  int chunks[2];
  long long chunk_rounds[2];
} rchol_amd_stats;

/* Level schedule of the two triangular sweeps (host analysis only). */
typedef struct rchol_amd_schedule {
  int nlevels[2];    /* [0] forward G y = r, [1] backward G^T z = y */
  int head_end[2];   /* levels [0, head_end) run in one workgroup */
  int tail_begin[2]; /* levels [tail_begin, nlevels) run in one workgroup */
  int mid_tasks[2];  /* 64-slot chunks of the levels in between (the engine
                        cuts narrow levels finer, by its resident waves) */
  int *level_ptr[2]; /* nlevels+1 slot offsets per level (malloc'ed) */
  int *slot_row[2];  /* n: row held by each slot, in level order (malloc'ed) */
  unsigned short *proxy[2]; /* n: per slot, the distance from its row's end of
                        the entry before the trailing entries of the level
                        just below, when that entry is at most 3 levels
                        below the row's; 0: none, and in levels of 2^18
                        rows or more (malloc'ed) */
} rchol_amd_schedule;

/* Thread-local text of the last error on the calling thread. */
const char *rchol_amd_last_error(void);

/* Context: one HIP device and one stream.  No reference counterpart (the
 * reference builds its MKL handles inside every pcg() call, pcg.cpp:22-27). */
int rchol_amd_create(int device, rchol_amd_ctx **ctx);
int rchol_amd_destroy(rchol_amd_ctx *ctx);

/* Device residency of A.  Replaces pcg::create_sparse for A
 * (c++/util/pcg.cpp:22, :31-54). */
int rchol_amd_upload_A(rchol_amd_ctx *ctx, size_t n, const size_t *rowPtr,
                       const size_t *colIdx, const double *val);

/* New values for the resident A, whose sparsity (rowPtr, colIdx) is
 * unchanged: one copy of val[nnz], no index work.  For a sequence of
 * matrices on one pattern, as python/rchol/ex_reuse_partition.py:28-42
 * factors with a reused ordering.  No reference counterpart (the
 * reference rebuilds its handles in every pcg() call, pcg.cpp:22-27);
 * nnz must equal the resident matrix's. */
int rchol_amd_update_A_values(rchol_amd_ctx *ctx, size_t nnz, const double *val);

/* Device residency of the factor, from the arrays rchol() stores
 * (rchol_lap.cpp:358-387): colPtr[n+1], rowIdx[nnz], val[nnz].  Builds the
 * forward (G) and backward (G^T) sweeps with their dependency-level
 * schedules.  Replaces pcg::create_sparse for G (c++/util/pcg.cpp:23,
 * :31-54) and the TRIANGULAR/UPPER descriptor of pcg.cpp:146-149. */
int rchol_amd_upload_G(rchol_amd_ctx *ctx, size_t n, const size_t *colPtr,
                       const size_t *rowIdx, const double *val, int precision);

/* z = G^-T G^-1 r on host buffers.  Replaces pcg::precond_solve
 * (c++/util/pcg.cpp:141-159). */
int rchol_amd_precond(rchol_amd_ctx *ctx, const double *r, double *z);

/* y = A x on host buffers.  Replaces pcg::matrix_vector_product
 * (c++/util/pcg.cpp:130-138). */
int rchol_amd_spmv(rchol_amd_ctx *ctx, const double *x, double *y);

/* PCG on host buffers.  Replaces pcg::iteration (c++/util/pcg.cpp:57-127):
 * r0 = b, stop when ||r||_2 <= tol*||b||_2 or after maxit iterations,
 * alpha = (p.r)/(p.q), beta = (r.z)/(r_prev.z_prev), relres = ||Ax-b||/||b||
 * on the true residual.  x (length n) is read as the starting accumulator
 * when x_in_is_initial != 0 (the reference keeps a pre-sized caller x,
 * pcg.cpp:67, while r still starts at b), else treated as zeros. */
int rchol_amd_pcg(rchol_amd_ctx *ctx, const double *b, double tol, int maxit, double *x,
                  int x_in_is_initial, double *relres, int *itr);

/* Same on device pointers (d_b, d_x on the context's device; d_x is
 * overwritten).  Used by the Python binding with torch tensors. */
int rchol_amd_pcg_device(rchol_amd_ctx *ctx, const double *d_b, double tol, int maxit,
                         double *d_x, int x_in_is_initial, double *relres, int *itr);

/* Preconditioner apply on device pointers. */
int rchol_amd_precond_device(rchol_amd_ctx *ctx, const double *d_r, double *d_z);

/* rchol_amd_precond_device without any device allocation: no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): the
 * sweep vectors, the permutation staging and the control block live in a
 * static 64 MiB device array that each GPU holds until the process ends.
 * Limit: y (and z for an FP32 factor) take 8 bytes per row, plus 16 per row
 * with a permutation set: about 8.4 M rows, or 2.8 M with a permutation;
 * larger factors get RCHOL_AMD_ERANGE.  Every call on one GPU shares that
 * GPU's array (it keeps the last call's contents; each call resets what it
 * reads), so calls that overlap on one GPU -- from several threads or
 * contexts -- race and must be serialised by the caller.  With an FP64
 * factor and no permutation, d_r and d_z must not overlap. */
int rchol_amd_precond_device_noalloc(rchol_amd_ctx *ctx, const double *d_r, double *d_z);

/* Batched PCG on k right-hand sides (1 <= k <= 32), stored interleaved:
 * B[i*k + c] is row i of column c (row-major n x k), X likewise.  Every
 * column runs the recurrence of rchol_amd_pcg with its own scalars and stops
 * on its own test; matrix entries are read once per row for all columns.
 * relres[k], itr[k] receive the per-column results.  No reference
 * counterpart: it generalises pcg::iteration (pcg.cpp:57-127) to a block
 * of independent right-hand sides. */
int rchol_amd_pcg_batch(rchol_amd_ctx *ctx, int k, const double *B, double tol, int maxit,
                        double *X, int x_in_is_initial, double *relres, int *itr);

/* Same on device pointers; k must be 4, 8, 16 or 32. */
int rchol_amd_pcg_batch_device(rchol_amd_ctx *ctx, int k, const double *d_B, double tol,
                               int maxit, double *d_X, int x_in_is_initial, double *relres,
                               int *itr);

/* Uneven batches: A and G are the block-diagonal stacks of nseg
 * independent systems (system s owns rows off[s] .. off[s+1]-1,
 * nseg <= 256).  rchol_amd_pcg_segments then runs one PCG per system
 * (pcg.cpp:57-127 semantics each, own scalars and stopping test) with every
 * stage of all systems in one launch; relres[nseg], itr[nseg].  No reference
 * counterpart (batched use of pcg::iteration). */
int rchol_amd_set_segments(rchol_amd_ctx *ctx, size_t nseg, const size_t *off);
int rchol_amd_pcg_segments(rchol_amd_ctx *ctx, const double *b, double tol, int maxit,
                           double *x, int x_in_is_initial, double *relres, int *itr);

/* Z = G^-T G^-1 R for k interleaved columns (pcg.cpp:141-159 per column). */
int rchol_amd_precond_batch(rchol_amd_ctx *ctx, int k, const double *R, double *Z);

/* One-shot solve with the exact argument list of the reference constructor
 * pcg(A, b, tol, maxit, G, x, relres, itr) (c++/util/pcg.hpp:15-16,
 * pcg.cpp:14-28): A as SparseCSR arrays, G as the stored factor arrays.
 * Uses device 0; creates and destroys a context per call. */
int rchol_amd_pcg_csr(size_t n, const size_t *aRowPtr, const size_t *aColIdx,
                      const double *aVal, const double *b, double tol, int maxit,
                      const size_t *gColPtr, const size_t *gRowIdx, const double *gVal,
                      double *x, int x_in_is_initial, double *relres, int *itr);

/* Declare the ordering P of the uploaded A and G (they factor A(P,P), as
 * returned by rchol(A, G, P, threads), rchol_parallel.hpp:6): from then on
 * the vectors passed to pcg / precond (single and batched) are in the
 * original numbering; the engine gathers b(P) (util.hpp:147-155) and
 * scatters x(P) = x' (ex_laplace.py:30-31) on the device.  perm == NULL
 * clears it. */
int rchol_amd_set_perm(rchol_amd_ctx *ctx, size_t n, const size_t *perm);

/* g[e] = -(lam[u[e]] - lam[v[e]]) * (x[u[e]] - x[v[e]]) on device pointers:
 * the gradient of a loss through x = A(w)^-1 b with respect to the weight of
 * edge (u, v) of a weighted Laplacian A(w), given lam = A^-1 dL/dx.  No
 * reference counterpart (used by the differentiable solve). */
int rchol_amd_edge_grad_device(rchol_amd_ctx *ctx, size_t m, const long long *d_u,
                               const long long *d_v, const double *d_lam, const double *d_x,
                               double *d_g);

/* Autotuner: times the preconditioner application with a minimum of 1, 2,
 * 4 and 8 lanes per sweep row and the SpMV in stream form and with 4, 8 and 16
 * lanes per row on
 * the uploaded matrices (median of reps runs) and keeps the fastest
 * (RCHOL_AMD_OPT_TRSV_LANES / _SPMV_LANES).  Outputs may be NULL.  No
 * reference counterpart. */
int rchol_amd_autotune(rchol_amd_ctx *ctx, int reps, int *trsv_lanes, int *spmv_lanes,
                       double *trsv_ms, double *spmv_ms);

/* Streaming device copy of `bytes` (median of reps launches): the achievable
 * HBM bandwidth (read + write bytes / time) and the known-size kernel used to
 * scale profiler byte counters.  No reference counterpart. */
int rchol_amd_copy_probe(rchol_amd_ctx *ctx, size_t bytes, int reps, double *gbs, double *ms);

int rchol_amd_set_option(rchol_amd_ctx *ctx, int key, double value);
int rchol_amd_get_stats(rchol_amd_ctx *ctx, rchol_amd_stats *out);

/* Host-only analysis of a stored factor (no device needed): the validation
 * and level schedule that rchol_amd_upload_G builds, for inspection and
 * tests.  Free with rchol_amd_schedule_free.  No reference counterpart. */
int rchol_amd_analyze_G(size_t n, const size_t *colPtr, const size_t *rowIdx, const double *val,
                        int serial_width, rchol_amd_schedule *out);
void rchol_amd_schedule_free(rchol_amd_schedule *s);

/* Host-only: the plan of the chunked sweeps (RCHOL_AMD_OPT_CHUNKED) for a
 * stored factor, chunks of at most max_rows rows, executed on the CPU in the
 * device kernel's order of operations: z = G^-T G^-1 r.  The plan's
 * invariants are checked on the way (a chunk reads only rows published
 * before it, a round reads only rows no later round writes); a violation is
 * RCHOL_AMD_EINVAL.  counts[3] / rounds / est_us[2] (may be NULL) receive
 * the chunks (both sweeps share them) and the external parts of the forward
 * and of the backward plan, both sweeps' rounds, and the plans'
 * critical-path estimates in microseconds.  For tests; no reference
 * counterpart. */
int rchol_amd_chunk_apply_host(size_t n, const size_t *colPtr, const size_t *rowIdx,
                               const double *val, int max_rows, const double *r, double *z,
                               int *counts, long long *rounds, double *est_us);

/* Number of visible HIP devices (0 when none); never fails. */
int rchol_amd_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* RCHOL_AMD_H */
