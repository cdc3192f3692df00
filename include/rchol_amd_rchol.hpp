// C++ drop-in for the reference's factorization entry points
//
//   void rchol(const SparseCSR &A, SparseCSR &G);                  c++/rchol/rchol.hpp:6
//   void rchol(const SparseCSR &A, SparseCSR &G,
//              std::vector<size_t> &permutation, int threads);     c++/rchol/rchol_parallel.hpp:6
//
// over librchol_factor.so (include/rchol_factor.h; link with -lrchol_factor).
// Same argument meaning and errors as the reference:
//  * threads must be positive and a power of two (std::invalid_argument with
//    the reference's texts, rchol_parallel.cpp:39-43); a graph too small for
//    the requested depth throws "too many threads requested"
//    (find_separator.cpp:123-126);
//  * the nested-dissection permutation is APPENDED to `permutation`
//    (rchol_parallel.cpp:62), G factors A(P, P) and is stored as rchol_lap
//    stores it: G.rowPtr/colIdx/val hold the CSC of lower G (column pointers,
//    row indices, values; rchol_lap.cpp:136-148, coalesce :358-387).
// Additions: an overload that also returns `part`, the block boundaries the
// reference computes internally (result_idx, rchol_parallel.cpp:63-69: 0,
// the prefix sums of the nested-dissection block sizes in post-order, the
// last one plus one for the Laplacian's extra vertex); and rchol_amd::
// factorize with an explicit sampling seed, ordering mode and, for
// grid-numbered matrices, the grid for geometric dissection.
//
// Separators come from rchol_nd_graph (breadth-first level structures)
// instead of METIS, so P differs from the reference's for the same A; the
// contract (post-order blocks, separator last) is the same.  The reference
// seeds every sampling thread from std::random_device; the overloads
// without a seed do the same, once per call.
//
// G's arrays are allocated with new[] (as rchol_lap's are); when the CSR
// type has an `ownMemory` member it is set, so the reference's SparseCSR
// releases them.  Any earlier arrays of G are not released (the reference
// overwrites them too).
#ifndef RCHOL_AMD_RCHOL_HPP
#define RCHOL_AMD_RCHOL_HPP

#include <algorithm>
#include <cstddef>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "rchol_factor.h"

namespace rchol_amd {

struct FactorOptions {
  int threads = 1;                  // leaves of the dissection (power of two)
  unsigned seed = 0;                // sampling seed (use_seed) ...
  bool use_seed = false;            // ... or std::random_device when false
  int order = RCHOL_ORDER_STATIC;   // or RCHOL_ORDER_MIS
  int grid[3] = {0, 0, 0};          // nx, ny, nz of a grid-numbered A (0: general graph)
};

namespace detail {

template <class T, class = void>
struct has_own_memory : std::false_type {};
template <class T>
struct has_own_memory<T, std::void_t<decltype(std::declval<T &>().ownMemory)>> : std::true_type {};

struct Freed {
  rchol_factor_out f{};
  ~Freed() {
    rchol_factor_free(f.colPtr);
    rchol_factor_free(f.rowIdx);
    rchol_factor_free(f.val);
    rchol_factor_free(f.perm);
  }
};

inline void check_threads(int threads) {
  if (threads <= 0) throw std::invalid_argument("thread number should be positive");
  if ((threads & (threads - 1)) != 0)
    throw std::invalid_argument("thread number should be a power of 2");
}

}  // namespace detail

// The general form: dissection with opt.threads leaves (none for 1, as
// rchol.cpp:7-36), factor, permutation appended to P, block boundaries
// (result_idx) into part.
template <class CSR>
void factorize(const CSR &A, CSR &G, std::vector<size_t> &P, std::vector<size_t> &part,
               const FactorOptions &opt) {
  detail::check_threads(opt.threads);
  const size_t n = A.N;
  if (n == 0) throw std::invalid_argument("empty matrix");
  int depth = 1;
  while ((1 << (depth - 1)) < opt.threads) depth++;
  std::vector<size_t> perm(n), blocks(((size_t)1 << depth) - 1);
  if (depth == 1) {
    for (size_t i = 0; i < n; i++) perm[i] = i;
    blocks[0] = n;
  } else {
    int rc;
    if (opt.grid[0] > 0) {
      if ((size_t)opt.grid[0] * opt.grid[1] * opt.grid[2] != n)
        throw std::invalid_argument("grid does not match the matrix size");
      rc = rchol_nd_grid(opt.grid[0], opt.grid[1], opt.grid[2], depth, perm.data(), blocks.data());
    } else {
      rc = rchol_nd_graph(n, A.rowPtr, A.colIdx, depth, perm.data(), blocks.data());
    }
    if (rc == -3) throw std::invalid_argument("too many threads requested");
    if (rc) throw std::runtime_error(std::string("rchol_amd: dissection failed: ") +
                                     rchol_factor_last_error());
  }
  const unsigned seed = opt.use_seed ? opt.seed : std::random_device{}();
  detail::Freed out;
  if (rchol_factor(n, A.rowPtr, A.colIdx, A.val, depth == 1 ? nullptr : perm.data(),
                   blocks.data(), blocks.size(), opt.order, seed, &out.f))
    throw std::runtime_error(std::string("rchol_amd: ") + rchol_factor_last_error());
  const size_t nnz = out.f.colPtr[n];
  G.N = n;
  G.rowPtr = new size_t[n + 1];
  G.colIdx = new size_t[nnz];
  G.val = new double[nnz];
  std::copy(out.f.colPtr, out.f.colPtr + n + 1, G.rowPtr);
  std::copy(out.f.rowIdx, out.f.rowIdx + nnz, G.colIdx);
  std::copy(out.f.val, out.f.val + nnz, G.val);
  if constexpr (detail::has_own_memory<CSR>::value) G.ownMemory = true;
  P.insert(P.end(), out.f.perm, out.f.perm + n);
  part.assign(1, 0);
  for (size_t b : blocks) part.push_back(part.back() + b);
  part.back()++;
}

}  // namespace rchol_amd

// rchol.hpp:6 -- sequential, natural order.
template <class CSR>
void rchol(const CSR &A, CSR &G) {
  std::vector<size_t> P, part;
  rchol_amd::factorize(A, G, P, part, rchol_amd::FactorOptions{});
}

// rchol_parallel.hpp:6 -- nested dissection with `threads` leaves.
template <class CSR>
void rchol(const CSR &A, CSR &G, std::vector<size_t> &permutation, int threads) {
  std::vector<size_t> part;
  rchol_amd::FactorOptions o;
  o.threads = threads;
  rchol_amd::factorize(A, G, permutation, part, o);
}

// As above, also returning the block boundaries (result_idx).
template <class CSR>
void rchol(const CSR &A, CSR &G, std::vector<size_t> &permutation, int threads,
           std::vector<size_t> &part) {
  rchol_amd::FactorOptions o;
  o.threads = threads;
  rchol_amd::factorize(A, G, permutation, part, o);
}

#endif  // RCHOL_AMD_RCHOL_HPP
