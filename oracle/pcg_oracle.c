/*
 * ORACLE -- test infrastructure only.  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load this code, and only as the checker.
 *
 * Plain-C FP64 restatement of the reference PCG solve path:
 *   oracle_pcg        <- c++/util/pcg.cpp:57-127  (pcg::iteration)
 *   oracle_precond    <- c++/util/pcg.cpp:141-159 (pcg::precond_solve)
 *   oracle_spmv       <- c++/util/pcg.cpp:130-138 (matrix_vector_product)
 * Matrices use the reference's SparseCSR arrays (64-bit indices,
 * c++/sparse.hpp:26-30).  The factor argument is exactly what rchol() stores
 * in its SparseCSR: the CSC arrays of the lower-triangular G, which read as
 * CSR are the upper triangle G^T (rchol_lap.cpp:358-387).  As in
 * pcg.cpp:146-155 the stored matrix is treated as UPPER/NON_UNIT and
 *   trsv(TRANSPOSE)      solves  G   y = r   (column sweep, forward),
 *   trsv(NON_TRANSPOSE)  solves  G^T z = y   (row sweep, backward).
 * MKL's internal summation order is not reproducible here (MKL is absent);
 * the triangular sweeps are sequential, SpMV and BLAS-1 use OpenMP when built
 * with it.  Pins: tests/test_oracle.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static void spmv(size_t n, const size_t *rp, const size_t *ci, const double *v,
                 const double *x, double *y) {
  long long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long long)n; i++) {
    double s = 0.0;
    for (size_t k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

static double dot(size_t n, const double *a, const double *b) {
  double s = 0.0;
  long long i;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (i = 0; i < (long long)n; i++) s += a[i] * b[i];
  return s;
}

static double nrm2(size_t n, const double *a) { return sqrt(dot(n, a, a)); }

static void axpy(size_t n, double alpha, const double *x, double *y) {
  long long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long long)n; i++) y[i] += alpha * x[i];
}

/* G y = r with G = (stored)^T: column sweep over the stored rows. */
static int fwd(size_t n, const size_t *sp, const size_t *si, const double *sv,
               const double *r, double *y) {
  memcpy(y, r, n * sizeof(double));
  for (size_t j = 0; j < n; j++) {
    double d = 0.0;
    for (size_t k = sp[j]; k < sp[j + 1]; k++)
      if (si[k] == j) d = sv[k];
    if (d == 0.0) return -1;
    double yj = y[j] / d;
    y[j] = yj;
    for (size_t k = sp[j]; k < sp[j + 1]; k++)
      if (si[k] != j) y[si[k]] -= sv[k] * yj;
  }
  return 0;
}

/* G^T z = y with G^T = stored (upper CSR): backward row sweep. */
static int bwd(size_t n, const size_t *sp, const size_t *si, const double *sv,
               const double *y, double *z) {
  for (size_t ii = n; ii-- > 0;) {
    double s = y[ii], d = 0.0;
    for (size_t k = sp[ii]; k < sp[ii + 1]; k++) {
      if (si[k] == ii)
        d = sv[k];
      else
        s -= sv[k] * z[si[k]];
    }
    if (d == 0.0) return -1;
    z[ii] = s / d;
  }
  return 0;
}

int oracle_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

int oracle_spmv(size_t n, const size_t *rp, const size_t *ci, const double *v,
                const double *x, double *y) {
  spmv(n, rp, ci, v, x, y);
  return 0;
}

int oracle_fwd(size_t n, const size_t *sp, const size_t *si, const double *sv,
               const double *r, double *y) {
  return fwd(n, sp, si, sv, r, y);
}

int oracle_bwd(size_t n, const size_t *sp, const size_t *si, const double *sv,
               const double *y, double *z) {
  return bwd(n, sp, si, sv, y, z);
}

/* pcg.cpp:141-159 */
int oracle_precond(size_t n, const size_t *sp, const size_t *si, const double *sv,
                   const double *r, double *z) {
  double *y = (double *)calloc(n ? n : 1, sizeof(double));
  if (!y) return -2;
  int rc = fwd(n, sp, si, sv, r, y);
  if (!rc) rc = bwd(n, sp, si, sv, y, z);
  free(y);
  return rc;
}

/*
 * pcg.cpp:57-127.  x holds the starting accumulator (the reference keeps the
 * caller's x after resize, pcg.cpp:67, while r starts at b, pcg.cpp:70-71).
 * rhist (optional, length maxit+1) receives ||r_k||_2 at each loop test.
 * Returns 0, or -1 on a zero pivot in G.
 */
int oracle_pcg(size_t n, const size_t *arp, const size_t *aci, const double *av,
               const double *b, double tol, int maxit, const size_t *gsp,
               const size_t *gsi, const double *gsv, double *x, double *relres,
               int *itr, double *rhist) {
  double *r = (double *)malloc((n ? n : 1) * sizeof(double));
  double *prev_r = (double *)calloc(n ? n : 1, sizeof(double));
  double *prev_cond = (double *)calloc(n ? n : 1, sizeof(double));
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  double *temp = (double *)calloc(n ? n : 1, sizeof(double));
  double *q = (double *)calloc(n ? n : 1, sizeof(double));
  int rc = 0;
  if (!r || !prev_r || !prev_cond || !p || !temp || !q) {
    rc = -2;
    goto done;
  }
  memcpy(r, b, n * sizeof(double));
  int k = 0;
  for (;;) {
    double rn = nrm2(n, r);
    if (rhist) rhist[k] = rn;
    if (!(rn > nrm2(n, b) * tol && k < maxit)) break;
    rc = oracle_precond(n, gsp, gsi, gsv, r, temp);
    if (rc) goto done;
    if (k == 0) {
      memcpy(p, temp, n * sizeof(double));
    } else {
      double d1 = dot(n, r, temp);
      double d2 = dot(n, prev_r, prev_cond);
      double beta = d1 / d2;
      long long i;
#pragma omp parallel for schedule(static)
      for (i = 0; i < (long long)n; i++) p[i] *= beta;
      axpy(n, 1.0, temp, p);
    }
    spmv(n, arp, aci, av, p, q);
    double d1 = dot(n, p, r);
    double d2 = dot(n, p, q);
    double alpha = d1 / d2;
    axpy(n, alpha, p, x);
    memcpy(prev_r, r, n * sizeof(double));
    memcpy(prev_cond, temp, n * sizeof(double));
    axpy(n, -alpha, q, r);
    k++;
  }
  spmv(n, arp, aci, av, x, q);
  axpy(n, -1.0, b, q);
  *relres = nrm2(n, q) / nrm2(n, b);
  *itr = k;
done:
  free(r);
  free(prev_r);
  free(prev_cond);
  free(p);
  free(temp);
  free(q);
  return rc;
}
