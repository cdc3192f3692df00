// The reference's C++ driver (c++/ex_laplace.cpp, ex_laplace_parallel.cpp)
// on this engine: 3D Poisson, seeded RHS, rchol factor, PCG through the
// drop-in pcg(...) constructor.  Usage: ex_laplace -n SIDE [-t LEAVES]
// [-tol TOL] [-maxit M] [-seed S] [-bfile B] [-xfile X]
// -bfile reads b (SIDE^3 raw FP64, original numbering) instead of drawing it;
// -xfile writes the solution x (raw FP64, original numbering).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "rchol_amd_pcg.hpp"
#include "rchol_factor.h"

// Layout of the reference SparseCSR (c++/sparse.hpp:26-30).
struct SparseCSR {
  size_t N = 0;
  size_t *rowPtr = nullptr;
  size_t *colIdx = nullptr;
  double *val = nullptr;
};

int main(int argc, char **argv) {
  int n = 3, threads = 1, maxit = 200;
  unsigned seed = 1;
  double tol = 1e-6;
  const char *bfile = nullptr, *xfile = nullptr;
  for (int i = 1; i + 1 < argc; i++) {
    if (!strcmp(argv[i], "-bfile")) bfile = argv[i + 1];
    if (!strcmp(argv[i], "-xfile")) xfile = argv[i + 1];
    if (!strcmp(argv[i], "-n")) n = atoi(argv[i + 1]);
    if (!strcmp(argv[i], "-t")) threads = atoi(argv[i + 1]);
    if (!strcmp(argv[i], "-tol")) tol = atof(argv[i + 1]);
    if (!strcmp(argv[i], "-maxit")) maxit = atoi(argv[i + 1]);
    if (!strcmp(argv[i], "-seed")) seed = (unsigned)atoi(argv[i + 1]);
  }
  SparseCSR A0;
  if (rchol_laplace_3d(n, &A0.N, &A0.rowPtr, &A0.colIdx, &A0.val)) return 1;
  const size_t N = A0.N;
  std::vector<double> b(N);
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> dis(-1.0, 1.0);
  for (auto &v : b) v = dis(gen);
  if (bfile) {
    FILE *fb = fopen(bfile, "rb");
    if (!fb || fread(b.data(), sizeof(double), N, fb) != N) {
      fprintf(stderr, "cannot read %zu doubles from %s\n", N, bfile);
      return 1;
    }
    fclose(fb);
  }
  // nested dissection with `threads` leaves (identity ordering for 1)
  int depth = 1;
  while ((1 << (depth - 1)) < threads) depth++;
  std::vector<size_t> perm(N), blocks((1u << depth) - 1);
  if (depth == 1) {
    for (size_t i = 0; i < N; i++) perm[i] = i;
    blocks[0] = N;
  } else if (rchol_nd_grid(n, n, n, depth, perm.data(), blocks.data())) {
    return 1;
  }
  rchol_factor_out f;
  if (rchol_factor(N, A0.rowPtr, A0.colIdx, A0.val, depth == 1 ? nullptr : perm.data(),
                   blocks.data(), blocks.size(), RCHOL_ORDER_STATIC, seed, &f)) {
    fprintf(stderr, "%s\n", rchol_factor_last_error());
    return 1;
  }
  SparseCSR G;
  G.N = f.n;
  G.rowPtr = f.colPtr;
  G.colIdx = f.rowIdx;
  G.val = f.val;
  // A(P,P) and b(P) as ex_laplace_parallel.cpp:38-39
  std::vector<size_t> inv(N);
  for (size_t i = 0; i < N; i++) inv[f.perm[i]] = i;
  std::vector<size_t> rp(N + 1, 0), ci;
  std::vector<double> av;
  std::vector<double> bp(N);
  for (size_t i = 0; i < N; i++) {
    const size_t o = f.perm[i];
    std::vector<std::pair<size_t, double>> row;
    for (size_t k = A0.rowPtr[o]; k < A0.rowPtr[o + 1]; k++)
      row.push_back({inv[A0.colIdx[k]], A0.val[k]});
    std::sort(row.begin(), row.end());
    for (auto &e : row) {
      ci.push_back(e.first);
      av.push_back(e.second);
    }
    rp[i + 1] = ci.size();
    bp[i] = b[o];
  }
  SparseCSR A;
  A.N = N;
  A.rowPtr = rp.data();
  A.colIdx = ci.data();
  A.val = av.data();
  printf("Fill-in ratio: %.3g\n", 2.0 * G.rowPtr[N] / A.rowPtr[N]);
  double relres = 0.0;
  int itr = 0;
  std::vector<double> x;
  try {
    pcg(A, bp, tol, maxit, G, x, relres, itr);
  } catch (const std::exception &e) {
    fprintf(stderr, "%s\n", e.what());
    return 2;
  }
  printf("# CG iterations: %d\n", itr);
  printf("Relative residual: %.3g\n", relres);
  if (xfile) {
    // x of A(P,P) back to the original numbering: x_orig[P[i]] = x[i]
    std::vector<double> xo(N);
    for (size_t i = 0; i < N; i++) xo[f.perm[i]] = x[i];
    FILE *fx = fopen(xfile, "wb");
    if (!fx || fwrite(xo.data(), sizeof(double), N, fx) != N) {
      fprintf(stderr, "cannot write %s\n", xfile);
      return 1;
    }
    fclose(fx);
  }
  rchol_factor_free(f.colPtr);
  rchol_factor_free(f.rowIdx);
  rchol_factor_free(f.val);
  rchol_factor_free(f.perm);
  rchol_factor_free(A0.rowPtr);
  rchol_factor_free(A0.colIdx);
  rchol_factor_free(A0.val);
  return 0;
}
