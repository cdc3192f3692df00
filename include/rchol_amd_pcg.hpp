// C++ drop-in for the reference's PCG entry point (c++/util/pcg.hpp:13-26):
//
//   pcg(const SparseCSR &A, const std::vector<double> &b, double tol, int maxit,
//       const SparseCSR &G, std::vector<double> &x, double &relres, int &itr);
//
// The constructor is the solve, as in the reference.  It works with the
// reference's SparseCSR (c++/sparse.hpp:10-31) or any type with the same
// members (N, rowPtr, colIdx, val); G is exactly what rchol() filled.
// Differences from the reference: errors throw std::runtime_error (the
// reference ignores MKL status codes), and the solve runs on GPU 0 through
// librchol_amd.so (link with -lrchol_amd).
//
// rchol_amd::Solver keeps A and G resident on the device across solves (the
// reference rebuilds its handles on every call, pcg.cpp:22-27).
#ifndef RCHOL_AMD_PCG_HPP
#define RCHOL_AMD_PCG_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "rchol_amd.h"

namespace rchol_amd {

inline void check(int rc) {
  if (rc != RCHOL_AMD_OK)
    throw std::runtime_error(std::string("rchol_amd: ") + rchol_amd_last_error());
}

class Solver {
 public:
  explicit Solver(int device = 0) { check(rchol_amd_create(device, &ctx_)); }
  ~Solver() { rchol_amd_destroy(ctx_); }
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;

  template <class CSR>
  void set_matrix(const CSR &A) {
    check(rchol_amd_upload_A(ctx_, A.N, A.rowPtr, A.colIdx, A.val));
  }
  template <class CSR>
  void set_factor(const CSR &G, int precision = RCHOL_AMD_FP64) {
    check(rchol_amd_upload_G(ctx_, G.N, G.rowPtr, G.colIdx, G.val, precision));
  }
  void set_option(int key, double value) { check(rchol_amd_set_option(ctx_, key, value)); }

  // pcg.cpp:57-127 semantics; x is resized to n and, as in the reference
  // (pcg.cpp:67-71), a caller x that already has n entries is kept as the
  // starting accumulator while r starts at b.
  void solve(const std::vector<double> &b, double tol, int maxit, std::vector<double> &x,
             double &relres, int &itr) {
    const bool keep = x.size() == b.size();
    x.resize(b.size());
    check(rchol_amd_pcg(ctx_, b.data(), tol, maxit, x.data(), keep ? 1 : 0, &relres, &itr));
  }

  void precond(const std::vector<double> &r, std::vector<double> &z) {
    z.resize(r.size());
    check(rchol_amd_precond(ctx_, r.data(), z.data()));
  }

  rchol_amd_stats stats() const {
    rchol_amd_stats s;
    check(rchol_amd_get_stats(ctx_, &s));
    return s;
  }

 private:
  rchol_amd_ctx *ctx_ = nullptr;
};

}  // namespace rchol_amd

// The reference call shape: a class whose constructor performs the solve.
class pcg {
 public:
  template <class CSR>
  pcg(const CSR &A, const std::vector<double> &b, double tol, int maxit, const CSR &G,
      std::vector<double> &x, double &relres, int &itr) {
    const bool keep = x.size() == b.size();
    x.resize(b.size());
    rchol_amd::check(rchol_amd_pcg_csr(A.N, A.rowPtr, A.colIdx, A.val, b.data(), tol, maxit,
                                       G.rowPtr, G.colIdx, G.val, x.data(), keep ? 1 : 0,
                                       &relres, &itr));
  }
};

#endif  // RCHOL_AMD_PCG_HPP
