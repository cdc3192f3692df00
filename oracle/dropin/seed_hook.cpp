// TEST INFRASTRUCTURE ONLY.  Linked into the reference's unchanged
// c++/ex_laplace.cpp driver (oracle/Makefile target `dropin`) so that its
// run is reproducible: every std::random_device draw returns the seed in
// the environment variable RCHOL_DROPIN_SEED (default 1).  The driver's
// right-hand side (util.hpp:50-54, one std::mt19937 seeded from
// random_device) and the factor's per-thread generators (rchol_lap.cpp
// :495, :596, :685) then start from that seed, and a test can rebuild b and
// G on the host and run the oracle on the same system.
#include <cstdlib>
#include <random>

std::random_device::result_type std::random_device::_M_getval() {
  const char *s = std::getenv("RCHOL_DROPIN_SEED");
  return s ? (std::random_device::result_type)std::strtoul(s, nullptr, 10) : 1u;
}
