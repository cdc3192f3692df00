// Exercises include/rchol_amd_rchol.hpp, the C++ rchol(...) entry points,
// for tests/test_rchol_cpp.py.  Usage: rchol_cpp_check SIDE OUTDIR
// Writes, per case, G (colptr/rowidx/val), P and part as raw arrays, and
// prints one line per argument check: "<threads> <exception text>".
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "rchol_amd_rchol.hpp"

// The reference SparseCSR's members (c++/sparse.hpp:26-30), owning its arrays.
struct SparseCSR {
  size_t N = 0;
  size_t *rowPtr = nullptr;
  size_t *colIdx = nullptr;
  double *val = nullptr;
  bool ownMemory = false;
  SparseCSR() = default;
  SparseCSR(const SparseCSR &) = delete;
  ~SparseCSR() {
    if (ownMemory) {
      delete[] rowPtr;
      delete[] colIdx;
      delete[] val;
    }
  }
};

template <class T>
static void dump(const std::string &path, const T *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f || fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("cannot write " + path);
  fclose(f);
}

static void dump_case(const std::string &dir, const char *tag, const SparseCSR &G,
                      const std::vector<size_t> &P, const std::vector<size_t> &part) {
  const std::string b = dir + "/" + tag;
  dump(b + "_cp.bin", G.rowPtr, G.N + 1);
  dump(b + "_ri.bin", G.colIdx, G.rowPtr[G.N]);
  dump(b + "_gv.bin", G.val, G.rowPtr[G.N]);
  dump(b + "_p.bin", P.data(), P.size());
  dump(b + "_part.bin", part.data(), part.size());
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const int side = atoi(argv[1]);
  const std::string dir = argv[2];
  SparseCSR A;
  A.ownMemory = false;  // malloc'ed by rchol_laplace_3d
  if (rchol_laplace_3d(side, &A.N, &A.rowPtr, &A.colIdx, &A.val)) return 1;
  {  // rchol.hpp:6 shape, seeded
    SparseCSR G;
    std::vector<size_t> P, part;
    rchol_amd::FactorOptions o;
    o.use_seed = true;
    o.seed = 5;
    rchol_amd::factorize(A, G, P, part, o);
    dump_case(dir, "seq5", G, P, part);
  }
  {  // general-graph dissection, 4 leaves, seeded
    SparseCSR G;
    std::vector<size_t> P, part;
    rchol_amd::FactorOptions o;
    o.threads = 4;
    o.use_seed = true;
    o.seed = 9;
    rchol_amd::factorize(A, G, P, part, o);
    dump_case(dir, "graph9", G, P, part);
  }
  {  // rchol_parallel.hpp:6 shape; P is appended to
    SparseCSR G;
    std::vector<size_t> P = {7, 7}, part;
    rchol(A, G, P, 8, part);
    dump_case(dir, "par8", G, P, part);
    SparseCSR G2;
    std::vector<size_t> P2;
    rchol(A, G2, P2, 2);
    SparseCSR G1;
    rchol(A, G1);
    printf("sizes %zu %zu %zu %zu\n", G2.N, P2.size(), G1.N, G1.rowPtr[G1.N]);
  }
  for (int t : {0, -4, 3, 6, 1 << 20}) {
    SparseCSR G;
    std::vector<size_t> P;
    try {
      rchol(A, G, P, t);
      printf("%d ok\n", t);
    } catch (const std::invalid_argument &e) {
      printf("%d %s\n", t, e.what());
    }
  }
  rchol_factor_free(A.rowPtr);
  rchol_factor_free(A.colIdx);
  rchol_factor_free(A.val);
  return 0;
}
