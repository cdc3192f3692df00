"""Writes the golden fixtures of tests/golden/*.npz.

Run in a container that has the reference sources and the oracle built
(``make -C oracle``): the factor G of every fixture comes from the
reference's own host factorization (oracle/_ref/librchol_ref.so, seeded),
and the expected preconditioner applications / PCG results from the oracle
(oracle/pcg_oracle.c).  The fixtures are data only: matrices, permutations,
seeds, right-hand sides and expected outputs.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def lap1d(n):
    e = np.ones(n)
    return sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], format="csr")


def varcoef3d(s, sigma, shift, seed):
    """7-point SDDM with face conductivities exp(N(0, sigma^2)), Dirichlet
    boundary faces, plus shift*I; index k + j*s + i*s^2 as laplace_3d.hpp."""
    rng = np.random.default_rng(seed)
    n = s ** 3
    idx = np.arange(n).reshape(s, s, s)  # [i, j, k] -> k + j s + i s^2
    rows, cols, vals = [], [], []
    diag = np.full(n, shift)
    for axis in range(3):
        # interior faces along axis
        a = np.take(idx, np.arange(s - 1), axis=axis).ravel()
        b = np.take(idx, np.arange(1, s), axis=axis).ravel()
        w = np.exp(rng.normal(0.0, sigma, a.size))
        rows += [a, b]
        cols += [b, a]
        vals += [-w, -w]
        np.add.at(diag, a, w)
        np.add.at(diag, b, w)
        # boundary faces (Dirichlet): conductivity to the outside
        for side in (0, s - 1):
            f = np.take(idx, [side], axis=axis).ravel()
            diag[f] += np.exp(rng.normal(0.0, sigma, f.size))
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


def nd_grid(s, depth):
    """Geometric nested dissection of an s^3 grid into 2^(depth-1) leaves:
    returns (perm, block sizes in post-order) with find_separator's layout
    (find_separator.cpp:144-150: [left p, right p, separator])."""
    def rec(box, d):
        (x0, x1), (y0, y1), (z0, z1) = box
        if d == depth:
            v = [k + j * s + i * s * s for i in range(z0, z1) for j in range(y0, y1)
                 for k in range(x0, x1)]
            return v, [len(v)]
        ext = [x1 - x0, y1 - y0, z1 - z0]
        ax = int(np.argmax(ext))
        lo, hi = box[ax]
        mid = (lo + hi) // 2
        L = list(box); L[ax] = (lo, mid)
        R = list(box); R[ax] = (mid + 1, hi)
        S = list(box); S[ax] = (mid, mid + 1)
        lp, lv = rec(tuple(L), d + 1)
        rp, rv = rec(tuple(R), d + 1)
        (a0, a1), (b0, b1), (c0, c1) = S
        sv = [k + j * s + i * s * s for i in range(c0, c1) for j in range(b0, b1)
              for k in range(a0, a1)]
        return lp + rp + sv, lv + rv + [len(sv)]
    p, v = rec(((0, s), (0, s), (0, s)), 1)
    return np.array(p, dtype=np.uint64), np.array(v, dtype=np.uint64)


def save(name, A, Gs, perm, blocks, seed, b, tol=1e-10, maxit=5000):
    z = O.precond(Gs, b)
    x, relres, itr, hist = O.pcg(A, b, tol, maxit, Gs, history=True)
    Ax = O.spmv(A, b)
    np.savez_compressed(
        os.path.join(OUT, name + ".npz"),
        a_indptr=A.indptr.astype(np.uint64), a_indices=A.indices.astype(np.uint64),
        a_data=A.data, g_colptr=Gs[0], g_rowidx=Gs[1], g_val=Gs[2],
        perm=np.asarray(perm, dtype=np.uint64), blocks=np.asarray(blocks, dtype=np.uint64),
        seed=np.array(seed), b=b, z=z, x=x, relres=np.array(relres), itr=np.array(itr),
        rhist=hist, Ab=Ax, tol=np.array(tol))
    print(f"{name}: n={A.shape[0]} nnzA={A.nnz} nnzG={Gs[1].size} itr={itr} relres={relres:.2e}")


def main():
    rng = np.random.default_rng(20261202)
    # 1D Dirichlet chain: the factor is the exact Cholesky factor.
    A = lap1d(16)
    Gs = O.ref_rchol_seq(A, 1)
    save("lap1d_16", A, Gs, np.arange(16), [17], 1, rng.uniform(-1, 1, 16))
    # reference laplace_3d, natural order, sequential rchol
    A = O.ref_laplace_3d(4)
    Gs = O.ref_rchol_seq(A, 11)
    save("lap3d_4_seq", A, Gs, np.arange(A.shape[0]), [A.shape[0] + 1], 11,
         rng.uniform(-1, 1, A.shape[0]))
    # reference laplace_3d, geometric ND with 4 leaves (depth 3), parallel path
    A = O.ref_laplace_3d(8)
    perm, blocks = nd_grid(8, 3)
    Gs = O.ref_rchol_nd(A, perm, blocks, 7)
    Ap = O.ref_reorder(A, perm)
    save("lap3d_8_nd4", Ap, Gs, perm, blocks, 7, rng.uniform(-1, 1, A.shape[0]))
    # variable-coefficient SDDM (sigma = 1.5, +1e-3 I), ND with 2 leaves
    A = varcoef3d(6, 1.5, 1e-3, 5)
    perm, blocks = nd_grid(6, 2)
    Gs = O.ref_rchol_nd(A, perm, blocks, 3)
    Ap = O.ref_reorder(A, perm)
    save("varcoef_6_nd2", Ap, Gs, perm, blocks, 3, rng.uniform(-1, 1, A.shape[0]))
    # reference laplace_3d(12), ND with 8 leaves (depth 4)
    A = O.ref_laplace_3d(12)
    perm, blocks = nd_grid(12, 4)
    Gs = O.ref_rchol_nd(A, perm, blocks, 2026)
    Ap = O.ref_reorder(A, perm)
    save("lap3d_12_nd8", Ap, Gs, perm, blocks, 2026, rng.uniform(-1, 1, A.shape[0]))


if __name__ == "__main__":
    main()
