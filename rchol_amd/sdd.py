"""SDD systems through the SDDM solve (adapter around the hot path).

The reference solves a symmetric diagonally dominant (SDD) system with
positive off-diagonals by the 2n x 2n SDDM extension
(matlab/rchol/sdd_to_sddm.m:2-17, used in matlab/ex_sdd.m:12-13,29):

    A = D + N + P  (N <= 0, P >= 0 off-diagonal parts)
    Ae = [[D + N, -P], [-P, D + N]],  be = [b; -b],  x = (xe[:n] - xe[n:]) / 2.

Ae is SDDM, so rchol + the device PCG apply unchanged.
"""
import numpy as np
import scipy.sparse as sp

from . import factor as F
from .solver import FP64, Solver


def sdd_3d(n):
    """The reference's SDD test matrix (matlab/rchol/sdd_3d.m:2-9): the 3D
    Laplacian with the sign of one direction's couplings flipped, + 4 I."""
    e = np.ones(n)
    I = sp.identity(n, format="csr")
    D = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], format="csr")
    A1 = sp.kron(D, I)
    A2 = sp.kron(I, D)
    A = sp.kron(A1, -I) + sp.kron(I, A1) + sp.kron(I, A2) + 4 * sp.identity(n ** 3)
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    A.sort_indices()
    return A


def sdd_to_sddm(A):
    """[[D + N, -P], [-P, D + N]] (sdd_to_sddm.m:2-17)."""
    A = sp.csr_matrix(A)
    Dg = sp.diags(A.diagonal())
    off = sp.csr_matrix(A - Dg)
    pos = off.multiply(off > 0)
    neg = off.multiply(off < 0)
    top = sp.hstack([Dg + neg, -pos])
    bot = sp.hstack([-pos, Dg + neg])
    Ae = sp.csr_matrix(sp.vstack([top, bot]))
    Ae.eliminate_zeros()
    Ae.sort_indices()
    return Ae


def extend_rhs(b):
    b = np.asarray(b, dtype=np.float64)
    return np.concatenate([b, -b])


def recover(xe):
    n = xe.size // 2
    return (xe[:n] - xe[n:]) / 2


def solve_sdd(A, b, tol=1e-10, maxit=5000, leaves=4, order="mis", seed=1, precision=FP64,
              device=0):
    """x with A x = b for SDD A: extension, nested dissection of the
    extended graph, rchol, device PCG.  Returns (x, relres_ext, itr)."""
    Ae = sdd_to_sddm(A)
    depth = int(np.log2(leaves)) + 1
    perm, blocks = F.nd_graph(Ae, depth)
    Gs, po, _ = F.factor(Ae, perm, blocks, order, seed)
    Ap = F.permute(Ae, po)
    with Solver(Ap, Gs, precision, device=device) as s:
        s.set_perm(po)
        xe, relres, itr = s.pcg(extend_rhs(b), tol, maxit)
    return recover(xe), relres, itr
