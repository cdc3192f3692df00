/*
 * rchol_factor -- host side of the solve engine: the randomized Cholesky
 * factorization that produces the factor G and the nested-dissection
 * ordering the GPU consumes, plus the problem generators used by the
 * benchmarks and tests.  Plain C ABI; arrays returned through out-pointers
 * are malloc'ed and released with rchol_factor_free.  Status: 0 success,
 * negative error (rchol_factor_last_error() has the text).
 *
 * Index arrays use the reference's size_t (c++/sparse.hpp:26-30).
 * A permutation p follows the reference convention xp[i] = x[p[i]]
 * (c++/util/util.hpp:147-155, util.cpp:16-56).
 */
#ifndef RCHOL_FACTOR_H
#define RCHOL_FACTOR_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ordering inside each nested-dissection block during factorization. */
#define RCHOL_ORDER_STATIC 0 /* the given permutation, as rchol_lap does */
#define RCHOL_ORDER_MIS 1    /* independent-set layers chosen during elimination */

typedef struct rchol_factor_out {
  size_t n;
  size_t *colPtr; /* n+1: CSC of lower G == the arrays rchol() stores in SparseCSR G */
  size_t *rowIdx; /* nnz */
  double *val;    /* nnz */
  size_t *perm;   /* n: final ordering (G factors A(perm, perm)) */
  double seconds; /* wall time of the elimination */
} rchol_factor_out;

const char *rchol_factor_last_error(void);
void rchol_factor_free(void *p);

/* Randomized Cholesky of an SDDM matrix A (CSR, original numbering).
 * perm (nullable = identity) and the block sizes `blocks` (post-order, as
 * find_separator's Separator_info::val, find_separator.cpp:138-150; nblocks
 * = 2^d - 1, or 1 for the sequential path) give the nested dissection.
 * The elimination of static mode restates rchol.cpp:7-36 (sequential) and
 * rchol_parallel.cpp:108-137 with rchol_lap.cpp:97-855 (threads over the ND
 * tree, per-thread std::mt19937 seeded with `seed`): for the same inputs and
 * a reference build whose std::random_device returns `seed`, G is
 * bit-identical.  MIS mode keeps the ND blocks and reorders vertices inside
 * each block by independent-set layers of the current elimination graph;
 * out->perm then differs from perm inside blocks. */
int rchol_factor(size_t n, const size_t *rowPtr, const size_t *colIdx, const double *val,
                 const size_t *perm, const size_t *blocks, size_t nblocks, int order,
                 unsigned int seed, rchol_factor_out *out);

/* Geometric nested dissection of an nx*ny*nz grid (index x + y*nx +
 * z*nx*ny) into 2^(depth-1) leaves: splits the longest extent at its middle
 * plane.  perm: n entries; blocks: 2^depth - 1 entries, post-order. */
int rchol_nd_grid(int nx, int ny, int nz, int depth, size_t *perm, size_t *blocks);

/* Nested dissection of a general graph with find_separator's contract
 * (find_separator.cpp:109-163): separators from breadth-first level
 * structures instead of METIS.  The graph is A without its diagonal. */
int rchol_nd_graph(size_t n, const size_t *rowPtr, const size_t *colIdx, int depth, size_t *perm,
                   size_t *blocks);

/* 7-point Dirichlet Laplacian on a side^3 grid (laplace_3d.hpp:8-65). */
int rchol_laplace_3d(int side, size_t *n, size_t **rowPtr, size_t **colIdx, double **val);

/* 7-point SDDM on an nx*ny*nz grid with face conductivities exp(g),
 * g ~ N(0, sigma^2) i.i.d. (seeded), Dirichlet boundary faces with their own
 * conductivities, plus shift*I.  Index x + y*nx + z*nx*ny; rows sorted.
 * An axis of extent 1 has no boundary faces (nz = 1 gives the 2D 5-point
 * operator). */
int rchol_varcoef_3d(int nx, int ny, int nz, double sigma, double shift, unsigned long long seed,
                     size_t *n, size_t **rowPtr, size_t **colIdx, double **val);

/* Weighted graph Laplacian of the 6-neighbour nx*ny*nz grid with the given
 * edge weights plus shift*I.  Edges are enumerated x-edges, then y, then z,
 * each in index order of their lower endpoint (rchol_grid_edges). */
int rchol_grid_laplacian(int nx, int ny, int nz, const double *w, double shift, size_t *n,
                         size_t **rowPtr, size_t **colIdx, double **val);

/* Endpoints of the 6-neighbour grid edges in rchol_grid_laplacian order:
 * m = 3*nx*ny*nz - nx*ny - ny*nz - nx*nz edges, u[e] < v[e]. */
int rchol_grid_edges(int nx, int ny, int nz, size_t *m, size_t **u, size_t **v);

/* B = A(perm, perm) with sorted rows: the reference's reorder(A, P, B)
 * (c++/util/util.cpp:16-56), threaded. */
int rchol_permute_csr(size_t n, const size_t *rowPtr, const size_t *colIdx, const double *val,
                      const size_t *perm, size_t **oRowPtr, size_t **oColIdx, double **oVal);

#ifdef __cplusplus
}
#endif

#endif /* RCHOL_FACTOR_H */
