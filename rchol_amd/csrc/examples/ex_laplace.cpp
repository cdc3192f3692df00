// The reference's C++ driver (c++/ex_laplace.cpp, ex_laplace_parallel.cpp)
// on this engine: 3D Poisson, seeded RHS, rchol factor, PCG through the
// drop-in rchol(...) and pcg(...) entry points.  Usage: ex_laplace -n SIDE [-t LEAVES]
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
#include "rchol_amd_rchol.hpp"

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
  // rchol(A, G, P, threads) as ex_laplace_parallel.cpp:29-32, with the
  // seed and the grid's geometric dissection made explicit
  SparseCSR G;
  std::vector<size_t> P, part;
  rchol_amd::FactorOptions opt;
  opt.threads = threads;
  opt.use_seed = true;
  opt.seed = seed;
  opt.grid[0] = opt.grid[1] = opt.grid[2] = n;
  try {
    rchol_amd::factorize(A0, G, P, part, opt);
  } catch (const std::exception &e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  // A(P,P) and b(P) as ex_laplace_parallel.cpp:38-39
  SparseCSR A;
  A.N = N;
  if (rchol_permute_csr(N, A0.rowPtr, A0.colIdx, A0.val, P.data(), &A.rowPtr, &A.colIdx,
                        &A.val))
    return 1;
  std::vector<double> bp(N);
  for (size_t i = 0; i < N; i++) bp[i] = b[P[i]];
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
    for (size_t i = 0; i < N; i++) xo[P[i]] = x[i];
    FILE *fx = fopen(xfile, "wb");
    if (!fx || fwrite(xo.data(), sizeof(double), N, fx) != N) {
      fprintf(stderr, "cannot write %s\n", xfile);
      return 1;
    }
    fclose(fx);
  }
  delete[] G.rowPtr;
  delete[] G.colIdx;
  delete[] G.val;
  rchol_factor_free(A.rowPtr);
  rchol_factor_free(A.colIdx);
  rchol_factor_free(A.val);
  rchol_factor_free(A0.rowPtr);
  rchol_factor_free(A0.colIdx);
  rchol_factor_free(A0.val);
  return 0;
}
