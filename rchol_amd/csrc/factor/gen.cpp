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
// A row's entry count follows from its position, so z-planes are filled in
// parallel straight into the returned arrays.
int assemble7(int nx, int ny, int nz, const std::vector<double> (&w)[3],
              const std::vector<double> &diag, size_t *n, size_t **rowPtr, size_t **colIdx,
              double **val) {
  const size_t N = (size_t)nx * ny * nz;
  const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
  auto count = [&](int x, int y, int z) {
    return (size_t)1 + (x > 0) + (x < nx - 1) + (y > 0) + (y < ny - 1) + (z > 0) + (z < nz - 1);
  };
  size_t *rp = (size_t *)malloc((N + 1) * sizeof(size_t));
  if (!rp) return -2;
  // entries before plane z: planes are alike except the first and the last
  std::vector<size_t> plane(nz + 1, 0);
  for (int z = 0; z < nz; z++) {
    size_t c = 0;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) c += count(x, y, z);
    plane[z + 1] = plane[z] + c;
  }
  const size_t nnz = plane[nz];
  size_t *ci = (size_t *)malloc(std::max<size_t>(nnz, 1) * sizeof(size_t));
  double *v = (double *)malloc(std::max<size_t>(nnz, 1) * sizeof(double));
  if (!ci || !v) {
    free(rp);
    free(ci);
    free(v);
    return -2;
  }
  rchol_host::parallel_for((size_t)nz, [&](size_t z0, size_t z1) {
    for (int z = (int)z0; z < (int)z1; z++) {
      size_t k = plane[z];
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          const size_t i = x + y * sy + z * sz;
          rp[i] = k;
          if (z > 0) { ci[k] = i - sz; v[k++] = -w[2][i - sz]; }
          if (y > 0) { ci[k] = i - sy; v[k++] = -w[1][i - sy]; }
          if (x > 0) { ci[k] = i - sx; v[k++] = -w[0][i - sx]; }
          ci[k] = i;
          v[k++] = diag[i];
          if (x < nx - 1) { ci[k] = i + sx; v[k++] = -w[0][i]; }
          if (y < ny - 1) { ci[k] = i + sy; v[k++] = -w[1][i]; }
          if (z < nz - 1) { ci[k] = i + sz; v[k++] = -w[2][i]; }
        }
    }
  }, 1);
  rp[N] = nnz;
  *n = N;
  *rowPtr = rp;
  *colIdx = ci;
  *val = v;
  return 0;
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
  // edges in rchol_grid_edges order: axis by axis, each in index order of
  // its lower endpoint; base[a] = the first edge of axis a
  const size_t mx = (size_t)(nx - 1) * ny * nz, my = (size_t)nx * (ny - 1) * nz;
  const size_t base[3] = {0, mx, mx + my};
  std::vector<double> w[3];
  for (int a = 0; a < 3; a++) w[a].assign(N, 0.0);
  std::vector<double> diag(N);
  rchol_host::parallel_for((size_t)nz, [&](size_t z0, size_t z1) {
    for (size_t z = z0; z < z1; z++)
      for (size_t y = 0; y < (size_t)ny; y++)
        for (size_t x = 0; x < (size_t)nx; x++) {
          const size_t i = x + y * nx + z * nx * ny;
          if (x < (size_t)nx - 1) w[0][i] = wt[base[0] + x + (nx - 1) * (y + ny * z)];
          if (y < (size_t)ny - 1) w[1][i] = wt[base[1] + x + nx * (y + (ny - 1) * z)];
          if (z < (size_t)nz - 1) w[2][i] = wt[base[2] + i];
        }
  }, 1);
  // each row sums its own faces (no writes into other rows), in the order a
  // sweep over the edges would add them: per axis the face below, then above
  const size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
  rchol_host::parallel_for((size_t)nz, [&](size_t z0, size_t z1) {
    for (size_t z = z0; z < z1; z++)
      for (size_t y = 0; y < (size_t)ny; y++)
        for (size_t x = 0; x < (size_t)nx; x++) {
          const size_t i = x + y * nx + z * nx * ny;
          const size_t c[3] = {x, y, z};
          double d = shift;
          for (int a = 0; a < 3; a++) {
            if (c[a] > 0) d += w[a][i - st[a]];
            d += w[a][i];
          }
          diag[i] = d;
        }
  }, 1);
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
