// Problem generators (host): the matrices of the benchmark and test
// configurations, as 64-bit CSR with sorted rows.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../../include/rchol_factor.h"
#include "parallel.hpp"

namespace {

int export_csr(const std::vector<size_t> &rp, const std::vector<size_t> &ci,
               const std::vector<double> &v, size_t *n, size_t **rowPtr, size_t **colIdx,
               double **val) {
  *n = rp.size() - 1;
  *rowPtr = (size_t *)malloc(rp.size() * sizeof(size_t));
  *colIdx = (size_t *)malloc(std::max<size_t>(ci.size(), 1) * sizeof(size_t));
  *val = (double *)malloc(std::max<size_t>(v.size(), 1) * sizeof(double));
  if (!*rowPtr || !*colIdx || !*val) return -2;
  memcpy(*rowPtr, rp.data(), rp.size() * sizeof(size_t));
  memcpy(*colIdx, ci.data(), ci.size() * sizeof(size_t));
  memcpy(*val, v.data(), v.size() * sizeof(double));
  return 0;
}

// 7-point stencil assembly: w[axis][lower endpoint] = conductivity of the
// face between idx and idx + stride(axis) (0 when outside); diag given.
int assemble7(int nx, int ny, int nz, const std::vector<double> (&w)[3],
              const std::vector<double> &diag, size_t *n, size_t **rowPtr, size_t **colIdx,
              double **val) {
  const size_t N = (size_t)nx * ny * nz;
  const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
  std::vector<size_t> rp, ci;
  std::vector<double> v;
  rp.reserve(N + 1);
  ci.reserve(7 * N);
  v.reserve(7 * N);
  rp.push_back(0);
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const size_t i = x + y * sy + z * sz;
        if (z > 0) { ci.push_back(i - sz); v.push_back(-w[2][i - sz]); }
        if (y > 0) { ci.push_back(i - sy); v.push_back(-w[1][i - sy]); }
        if (x > 0) { ci.push_back(i - sx); v.push_back(-w[0][i - sx]); }
        ci.push_back(i);
        v.push_back(diag[i]);
        if (x < nx - 1) { ci.push_back(i + sx); v.push_back(-w[0][i]); }
        if (y < ny - 1) { ci.push_back(i + sy); v.push_back(-w[1][i]); }
        if (z < nz - 1) { ci.push_back(i + sz); v.push_back(-w[2][i]); }
        rp.push_back(ci.size());
      }
  return export_csr(rp, ci, v, n, rowPtr, colIdx, val);
}

}  // namespace

extern "C" {

// laplace_3d.hpp:8-65: index k + j n + i n^2, entries in the order
// -n^2, -n, -1, self (6), +1, +n, +n^2 (ascending columns).
int rchol_laplace_3d(int side, size_t *n, size_t **rowPtr, size_t **colIdx, double **val) {
  if (side < 1 || !n || !rowPtr || !colIdx || !val) return -1;
  const size_t N = (size_t)side * side * side;
  std::vector<double> w[3];
  for (int a = 0; a < 3; a++) w[a].assign(N, 1.0);
  std::vector<double> diag(N, 6.0);
  return assemble7(side, side, side, w, diag, n, rowPtr, colIdx, val);
}

int rchol_varcoef_3d(int nx, int ny, int nz, double sigma, double shift, unsigned long long seed,
                     size_t *n, size_t **rowPtr, size_t **colIdx, double **val) {
  if (nx < 1 || ny < 1 || nz < 1 || !n || !rowPtr || !colIdx || !val) return -1;
  const size_t N = (size_t)nx * ny * nz;
  const size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
  const int ext[3] = {nx, ny, nz};
  std::mt19937_64 gen(seed);
  std::normal_distribution<double> g(0.0, sigma);
  std::vector<double> w[3];
  std::vector<double> diag(N, shift);
  for (int a = 0; a < 3; a++) {
    w[a].assign(N, 0.0);
    for (size_t i = 0; i < N; i++) {
      const int c = (int)((i / st[a]) % ext[a]);
      if (c < ext[a] - 1) {
        const double k = std::exp(g(gen));
        w[a][i] = k;
        diag[i] += k;
        diag[i + st[a]] += k;
      }
    }
  }
  // Dirichlet boundary faces: conductivity to the outside value 0 (an axis
  // of extent 1 is a collapsed dimension -- 2D problems -- and has none)
  for (int a = 0; a < 3; a++) {
    if (ext[a] == 1) continue;
    for (size_t i = 0; i < N; i++) {
      const int c = (int)((i / st[a]) % ext[a]);
      if (c == 0) diag[i] += std::exp(g(gen));
      if (c == ext[a] - 1) diag[i] += std::exp(g(gen));
    }
  }
  return assemble7(nx, ny, nz, w, diag, n, rowPtr, colIdx, val);
}

int rchol_grid_edges(int nx, int ny, int nz, size_t *m, size_t **u, size_t **v) {
  if (nx < 1 || ny < 1 || nz < 1 || !m || !u || !v) return -1;
  const size_t N = (size_t)nx * ny * nz;
  const size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
  const int ext[3] = {nx, ny, nz};
  std::vector<size_t> U, V;
  for (int a = 0; a < 3; a++)
    for (size_t i = 0; i < N; i++)
      if ((int)((i / st[a]) % ext[a]) < ext[a] - 1) {
        U.push_back(i);
        V.push_back(i + st[a]);
      }
  *m = U.size();
  *u = (size_t *)malloc(std::max<size_t>(U.size(), 1) * sizeof(size_t));
  *v = (size_t *)malloc(std::max<size_t>(V.size(), 1) * sizeof(size_t));
  if (!*u || !*v) return -2;
  memcpy(*u, U.data(), U.size() * sizeof(size_t));
  memcpy(*v, V.data(), V.size() * sizeof(size_t));
  return 0;
}

int rchol_grid_laplacian(int nx, int ny, int nz, const double *wt, double shift, size_t *n,
                         size_t **rowPtr, size_t **colIdx, double **val) {
  if (nx < 1 || ny < 1 || nz < 1 || !wt || !n || !rowPtr || !colIdx || !val) return -1;
  const size_t N = (size_t)nx * ny * nz;
  const size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
  const int ext[3] = {nx, ny, nz};
  std::vector<double> w[3];
  std::vector<double> diag(N, shift);
  size_t e = 0;
  for (int a = 0; a < 3; a++) {
    w[a].assign(N, 0.0);
    for (size_t i = 0; i < N; i++)
      if ((int)((i / st[a]) % ext[a]) < ext[a] - 1) {
        const double k = wt[e++];
        w[a][i] = k;
        diag[i] += k;
        diag[i + st[a]] += k;
      }
  }
  return assemble7(nx, ny, nz, w, diag, n, rowPtr, colIdx, val);
}

// util.cpp:16-56: B = A(P, P), row i of B is row P[i] of A with columns
// renumbered by the inverse permutation and sorted (threaded).
int rchol_permute_csr(size_t n, const size_t *rowPtr, const size_t *colIdx, const double *val,
                      const size_t *perm, size_t **oRowPtr, size_t **oColIdx, double **oVal) {
  if (!rowPtr || !colIdx || !val || !perm || !oRowPtr || !oColIdx || !oVal) return -1;
  std::vector<size_t> inv(n, (size_t)-1);
  for (size_t i = 0; i < n; i++) {
    if (perm[i] >= n || inv[perm[i]] != (size_t)-1) return -1;
    inv[perm[i]] = i;
  }
  const size_t nnz = rowPtr[n];
  size_t *rp = (size_t *)malloc((n + 1) * sizeof(size_t));
  size_t *ci = (size_t *)malloc(std::max<size_t>(nnz, 1) * sizeof(size_t));
  double *v = (double *)malloc(std::max<size_t>(nnz, 1) * sizeof(double));
  if (!rp || !ci || !v) return -2;
  rp[0] = 0;
  for (size_t i = 0; i < n; i++) rp[i + 1] = rp[i] + (rowPtr[perm[i] + 1] - rowPtr[perm[i]]);
  rchol_host::parallel_for(n, [&](size_t i0, size_t i1) {
    std::vector<std::pair<size_t, double>> row;
    for (size_t i = i0; i < i1; i++) {
      const size_t o = perm[i];
      row.clear();
      for (size_t k = rowPtr[o]; k < rowPtr[o + 1]; k++) row.push_back({inv[colIdx[k]], val[k]});
      std::sort(row.begin(), row.end(),
                [](const std::pair<size_t, double> &a, const std::pair<size_t, double> &b) {
                  return a.first < b.first;
                });
      size_t at = rp[i];
      for (auto &e : row) {
        ci[at] = e.first;
        v[at] = e.second;
        at++;
      }
    }
  });
  *oRowPtr = rp;
  *oColIdx = ci;
  *oVal = v;
  return 0;
}

}  // extern "C"
