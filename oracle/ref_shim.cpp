// TEST INFRASTRUCTURE ONLY.  C entry points over the reference's own host
// factorization, compiled from the sources where they lie (see
// oracle/Makefile, target `ref`).  Nothing in the product links this file;
// only tests/ and the golden-vector script load oracle/_ref/librchol_ref.so.
//
// Two things are added around the reference code, neither of which replaces
// any of its functions:
//  * a seed hook: every std::random_device draw returns `g_seed`, so each
//    thread-local std::mt19937 in rchol_lap.cpp (:495, :596, :685) starts from
//    a known state and the factor becomes reproducible;
//  * each call runs on a fresh std::thread, so (a) the calling thread's
//    thread_local generator never carries state from an earlier call and
//    (b) the sched_setaffinity pinning in rchol_lap.cpp:100-114 does not
//    stick to the caller.
// stdout is redirected to /dev/null for the duration of a call because the
// reference prints per-leaf timings.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>
#include <unistd.h>
#include <fcntl.h>

#include "sparse.hpp"
#include "rchol/rchol.hpp"
#include "rchol_lap/rchol_lap.hpp"
#include "util/util.hpp"
#include "util/laplace_3d.hpp"

static unsigned int g_seed = 0;

// Seed hook (resolved inside this .so thanks to -Bsymbolic-functions).
std::random_device::result_type std::random_device::_M_getval() { return g_seed; }

namespace {

struct QuietStdout {
  int saved = -1;
  QuietStdout() {
    fflush(stdout);
    saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (devnull >= 0) { dup2(devnull, 1); close(devnull); }
  }
  ~QuietStdout() {
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
  }
};

void export_factor(SparseCSR &G, size_t **colPtr, size_t **rowIdx, double **val, size_t *nnz) {
  size_t n = G.N;
  size_t z = G.rowPtr[n];
  *colPtr = (size_t *)malloc((n + 1) * sizeof(size_t));
  *rowIdx = (size_t *)malloc((z ? z : 1) * sizeof(size_t));
  *val = (double *)malloc((z ? z : 1) * sizeof(double));
  memcpy(*colPtr, G.rowPtr, (n + 1) * sizeof(size_t));
  memcpy(*rowIdx, G.colIdx, z * sizeof(size_t));
  memcpy(*val, G.val, z * sizeof(double));
  *nnz = z;
}

}  // namespace

extern "C" {

// rchol.cpp:7-36 (sequential path, natural order).
int ref_rchol_seq(size_t n, const size_t *rowPtr, const size_t *colIdx, const double *val,
                  unsigned int seed, size_t **gColPtr, size_t **gRowIdx, double **gVal,
                  size_t *nnzG) {
  std::vector<size_t> rp(rowPtr, rowPtr + n + 1), ci(colIdx, colIdx + rowPtr[n]);
  std::vector<double> v(val, val + rowPtr[n]);
  int rc = 0;
  std::thread t([&]() {
    QuietStdout q;
    g_seed = seed;
    SparseCSR A(rp, ci, v, true);
    SparseCSR G;
    rchol(A, G);
    // rchol_lap hands over new[]-allocated arrays; G does not own them.
    export_factor(G, gColPtr, gRowIdx, gVal, nnzG);
    delete[] G.rowPtr;
    delete[] G.colIdx;
    delete[] G.val;
    G.N = 0;
  });
  t.join();
  return rc;
}

// rchol_parallel.cpp:108-137 with the ND permutation and block sizes supplied
// by the caller instead of find_separator (METIS is not available).  `val`
// holds the block sizes in post-order, exactly as Separator_info::val.
int ref_rchol_nd(size_t n, const size_t *rowPtr, const size_t *colIdx, const double *val,
                 const size_t *perm, const size_t *blk, size_t nblk, unsigned int seed,
                 size_t **gColPtr, size_t **gRowIdx, double **gVal, size_t *nnzG) {
  std::vector<size_t> rp(rowPtr, rowPtr + n + 1), ci(colIdx, colIdx + rowPtr[n]);
  std::vector<double> v(val, val + rowPtr[n]);
  std::vector<size_t> permutation(perm, perm + n);
  int rc = 0;
  std::thread t([&]() {
    QuietStdout q;
    g_seed = seed;
    SparseCSR A(rp, ci, v, true);
    std::vector<size_t> result_idx;
    result_idx.push_back(0);
    for (size_t i = 0; i < nblk; i++) result_idx.push_back(blk[i] + result_idx.back());
    result_idx[result_idx.size() - 1]++;
    std::vector<size_t> rowPtrP, colIdxP;
    std::vector<double> valP;
    reorder(A, rowPtrP, colIdxP, valP, permutation);
    std::vector<size_t> rowPtrU, colIdxU;
    std::vector<double> valU;
    triu_csr(rowPtrP, colIdxP, valP, rowPtrU, colIdxU, valU);
    std::vector<size_t> rowPtrL, colIdxL;
    std::vector<double> valL;
    convert_to_laplace(rowPtrU, colIdxU, valU, rowPtrL, colIdxL, valL, rowPtrP, colIdxP, valP);
    change_to_positive_edge(valL);
    SparseCSR G;
    rchol_lap(rowPtrL, colIdxL, valL, G.rowPtr, G.colIdx, G.val, G.N, result_idx);
    export_factor(G, gColPtr, gRowIdx, gVal, nnzG);
    delete[] G.rowPtr;
    delete[] G.colIdx;
    delete[] G.val;
    G.N = 0;
  });
  t.join();
  return rc;
}

// laplace_3d.hpp:8-65 via util.cpp:6-12.
int ref_laplace_3d(int side, size_t *n, size_t **rowPtr, size_t **colIdx, double **val) {
  std::vector<size_t> rp, ci;
  std::vector<double> v;
  laplace_3d<double>((size_t)side, rp, ci, v);
  *n = rp.size() - 1;
  *rowPtr = (size_t *)malloc(rp.size() * sizeof(size_t));
  *colIdx = (size_t *)malloc(ci.size() * sizeof(size_t));
  *val = (double *)malloc(v.size() * sizeof(double));
  memcpy(*rowPtr, rp.data(), rp.size() * sizeof(size_t));
  memcpy(*colIdx, ci.data(), ci.size() * sizeof(size_t));
  memcpy(*val, v.data(), v.size() * sizeof(double));
  return 0;
}

// util.cpp:16-56: B = A(P,P) with per-row sorted columns.
int ref_reorder(size_t n, const size_t *rowPtr, const size_t *colIdx, const double *val,
                const size_t *perm, size_t **oRowPtr, size_t **oColIdx, double **oVal) {
  std::vector<size_t> rp(rowPtr, rowPtr + n + 1), ci(colIdx, colIdx + rowPtr[n]);
  std::vector<double> v(val, val + rowPtr[n]);
  std::vector<size_t> permutation(perm, perm + n);
  SparseCSR A(rp, ci, v, true);
  std::vector<size_t> rpo, cio;
  std::vector<double> vo;
  reorder(A, rpo, cio, vo, permutation);
  *oRowPtr = (size_t *)malloc(rpo.size() * sizeof(size_t));
  *oColIdx = (size_t *)malloc((cio.size() ? cio.size() : 1) * sizeof(size_t));
  *oVal = (double *)malloc((vo.size() ? vo.size() : 1) * sizeof(double));
  memcpy(*oRowPtr, rpo.data(), rpo.size() * sizeof(size_t));
  memcpy(*oColIdx, cio.data(), cio.size() * sizeof(size_t));
  memcpy(*oVal, vo.data(), vo.size() * sizeof(double));
  return 0;
}

void ref_free(void *p) { free(p); }

}  // extern "C"
