// TEST INFRASTRUCTURE ONLY.  A forwarding header named like the reference's
// c++/util/pcg.hpp: placed ahead of c++/util on the include path, it makes
// the reference's unchanged drivers (c++/ex_laplace.cpp:8, :42) call this
// engine's drop-in pcg(...) constructor instead of the MKL one.  That is the
// whole integration on the caller's side (plus -lrchol_amd at link time).
#ifndef RCHOL_DROPIN_PCG_HPP
#define RCHOL_DROPIN_PCG_HPP
#include "rchol_amd_pcg.hpp"
#endif
