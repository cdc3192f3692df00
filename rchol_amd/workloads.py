"""The benchmark workloads, built on the host with this package's own
generator, nested dissection and factorization (librchol_factor.so).

  laplace_natural(side)              32^3 7-point Laplacian (laplace_3d.hpp:8-65),
                                     sequential rchol in natural order
                                     (ex_laplace.cpp:22-31)
  laplace_nd(side, leaves, order)    7-point Laplacian, nested dissection with
                                     `leaves` leaves (ex_laplace_parallel.cpp:25-39)
  varcoef_nd(side, leaves, order)    variable-coefficient 7-point Poisson, face
                                     conductivities exp(N(0, 1.5^2)) + 1e-3 I
  uneven_2d(nsys)                    block of independent 2D 5-point SDDM systems,
                                     sides log-uniform in [64, 1024], log-normal
                                     edge weights (sigma 1)

Each returns the permuted matrix A(perm, perm) as scipy CSR (what the
reference's reorder() hands to pcg(), util.cpp:16-56), the stored factor
triple exactly as rchol() fills SparseCSR G (rchol_lap.cpp:358-387) and the
permutation.  order="static" keeps the reference's order inside the ND
blocks (find_separator.cpp:111-121); order="mis" numbers each block by
independent-set layers while it is eliminated.
"""
import numpy as np

from . import factor as F

SEED_A = 20261202   # conductivity field of the variable-coefficient problem
SEED_G = 1          # sampling seed of the factorization
SEED_B = 7          # right-hand side


def _depth(leaves):
    d = int(np.log2(leaves))
    if 1 << d != leaves:
        raise ValueError("leaves must be a power of two")
    return d + 1


def rhs(n, seed=SEED_B):
    """b ~ U[-1, 1], seeded."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


def laplace_natural(side=32, seed=SEED_G):
    A = F.laplace_3d(side)
    n = A.shape[0]
    Gs, po, _ = F.factor(A, None, [n], "static", seed)
    return A, Gs, po


def laplace_nd(side=128, leaves=64, order="mis", seed=SEED_G):
    A = F.laplace_3d(side)
    perm, blocks = F.nd_grid(side, side, side, _depth(leaves))
    Gs, po, _ = F.factor(A, perm, blocks, order, seed)
    return F.permute(A, po), Gs, po


def varcoef_nd(side=256, leaves=128, order="mis", seed_a=SEED_A, seed=SEED_G):
    A = F.varcoef_3d(side, side, side, 1.5, 1e-3, seed_a)
    perm, blocks = F.nd_grid(side, side, side, _depth(leaves))
    Gs, po, _ = F.factor(A, perm, blocks, order, seed)
    Ap = F.permute(A, po)
    return Ap, Gs, po


def uneven_sides(nsys=48, seed=48):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(64), np.log(1024), nsys)).astype(int), rng


def uneven_2d(nsys=48, seed=48, leaves=8, order="mis"):
    """[(A_s(perm, perm), Gs_s)] and the right-hand sides [b_s] (permuted
    numbering) of nsys independent 2D 5-point SDDM systems."""
    sides, rng = uneven_sides(nsys, seed)
    systems, bs = [], []
    for i, sd in enumerate(sides):
        sd = int(sd)
        A = F.varcoef_3d(sd, sd, 1, 1.0, 0.0, 1000 + i)
        perm, blocks = F.nd_grid(sd, sd, 1, _depth(leaves))
        Gs, po, _ = F.factor(A, perm, blocks, order, i + 1)
        systems.append((F.permute(A, po), Gs))
        bs.append(rng.uniform(-1, 1, A.shape[0]))
    return systems, bs, sides
