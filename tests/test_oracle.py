"""Pins of the oracle (oracle/pcg_oracle.c, oracle/_ref) -- CPU only.

Each check is chosen so that a plausible mistake in the oracle (wrong sweep
orientation, transposed factor, wrong alpha/beta, dropped term) fails it:
library routines (scipy), closed forms, invariants and brute force.
"""
import glob
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    n = d["a_indptr"].size - 1
    A = sp.csr_matrix((d["a_data"], d["a_indices"], d["a_indptr"]), shape=(n, n))
    Gs = (d["g_colptr"], d["g_rowidx"], d["g_val"])
    return A, Gs, d


FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz")))


def lap1d(n):
    e = np.ones(n)
    return sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], format="csr")


def random_lower(n, density, seed):
    rng = np.random.default_rng(seed)
    L = sp.random(n, n, density=density, random_state=rng, format="csr")
    L = sp.tril(L, -1) * -0.3 + sp.diags(rng.uniform(1.0, 2.0, n))
    return sp.csr_matrix(L)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_precond_matches_scipy_triangular_solves(seed):
    G = random_lower(60, 0.1, seed)
    Gs = (sp.csc_matrix(G).indptr, sp.csc_matrix(G).indices, sp.csc_matrix(G).data)
    r = np.random.default_rng(seed).standard_normal(60)
    y = spl.spsolve_triangular(G.tocsr(), r, lower=True)
    z = spl.spsolve_triangular(G.T.tocsr(), y, lower=False)
    np.testing.assert_allclose(O.fwd(Gs, r), y, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.precond(Gs, r), z, rtol=1e-12, atol=1e-12)
    # dense brute force: z = (G G^T)^-1 r
    Gd = G.toarray()
    np.testing.assert_allclose(O.precond(Gs, r), np.linalg.solve(Gd @ Gd.T, r), rtol=1e-10)


def test_precond_orientation_is_not_transposed():
    # G^-T G^-1 differs from G^-1 G^-T for a non-symmetric structure: the
    # oracle must pick the first (pcg.cpp:151 TRANSPOSE then :155).
    G = random_lower(30, 0.2, 5).toarray()
    Gs_c = sp.csc_matrix(G)
    Gs = (Gs_c.indptr, Gs_c.indices, Gs_c.data)
    r = np.arange(1.0, 31.0)
    want = sla.solve_triangular(G.T, sla.solve_triangular(G, r, lower=True), lower=False)
    wrong = sla.solve_triangular(G, sla.solve_triangular(G.T, r, lower=False), lower=True)
    got = O.precond(Gs, r)
    assert np.allclose(got, want, rtol=1e-12)
    assert not np.allclose(got, wrong, rtol=1e-6)


def test_spmv_matches_scipy():
    A = O.ref_laplace_3d(5) if O.ref_available() else lap1d(125)
    x = np.random.default_rng(3).standard_normal(A.shape[0])
    np.testing.assert_allclose(O.spmv(A, x), A @ x, rtol=1e-14, atol=1e-14)


def test_pcg_exact_factor_converges_in_one_iteration():
    # Exact Cholesky G: M^-1 = A^-1, so one PCG step solves the system.
    A = lap1d(40)
    Gd = np.linalg.cholesky(A.toarray())
    Gc = sp.csc_matrix(Gd)
    b = np.random.default_rng(0).uniform(-1, 1, 40)
    x, relres, itr = O.pcg(A, b, 1e-10, 100, (Gc.indptr, Gc.indices, Gc.data))
    assert itr == 1
    np.testing.assert_allclose(x, np.linalg.solve(A.toarray(), b), rtol=1e-10)
    assert relres < 1e-13


def test_pcg_iterates_match_textbook_jacobi_pcg():
    # G = diag(sqrt(A_ii)) gives Jacobi PCG; scipy's cg with M = D^-1 follows
    # the same Krylov iterates.  A wrong alpha or beta formula changes x_k.
    A = O.ref_laplace_3d(6) if O.ref_available() else lap1d(216)
    A = sp.csr_matrix(A + sp.diags(np.linspace(0.0, 3.0, A.shape[0])))
    n = A.shape[0]
    d = A.diagonal()
    Gs = (np.arange(n + 1), np.arange(n), np.sqrt(d))
    b = np.random.default_rng(1).uniform(-1, 1, n)
    ref_iter = []
    M = sp.diags(1.0 / d)
    spl.cg(A, b, rtol=1e-30, atol=0.0, maxiter=6, M=M, callback=lambda xk: ref_iter.append(xk.copy()))
    for k in range(1, 7):
        xk, _, itr = O.pcg(A, b, 1e-30, k, Gs)
        assert itr == k
        np.testing.assert_allclose(xk, ref_iter[k - 1], rtol=1e-9, atol=1e-12)


def test_pcg_solution_and_true_residual():
    A = lap1d(50) + sp.identity(50) * 0.01
    Gd = np.diag(np.sqrt(A.diagonal()))
    Gc = sp.csc_matrix(Gd)
    b = np.random.default_rng(2).uniform(-1, 1, 50)
    x, relres, itr, hist = O.pcg(A, b, 1e-12, 500, (Gc.indptr, Gc.indices, Gc.data),
                                 history=True)
    np.testing.assert_allclose(x, spl.spsolve(sp.csc_matrix(A), b), rtol=1e-9)
    assert relres == pytest.approx(np.linalg.norm(A @ x - b) / np.linalg.norm(b), rel=1e-12)
    # stopping rule of pcg.cpp:82 on the recursive residual
    assert hist[-1] <= 1e-12 * np.linalg.norm(b) and hist[-2] > 1e-12 * np.linalg.norm(b)


def test_pcg_maxit_and_initial_x_semantics():
    A = lap1d(30)
    Gc = sp.csc_matrix(np.diag(np.sqrt(A.diagonal())))
    Gs = (Gc.indptr, Gc.indices, Gc.data)
    b = np.ones(30)
    _, _, itr = O.pcg(A, b, 1e-14, 3, Gs)
    assert itr == 3
    # pcg.cpp:67-71: a pre-sized x is kept as accumulator while r starts at b
    x0 = np.full(30, 0.5)
    x1, _, _ = O.pcg(A, b, 1e-14, 4, Gs, x0=x0)
    x2, _, _ = O.pcg(A, b, 1e-14, 4, Gs)
    np.testing.assert_allclose(x1 - x0, x2, rtol=1e-13, atol=1e-14)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_reference_laplace_3d_closed_form():
    for s in (3, 5):
        A = O.ref_laplace_3d(s)
        n = s ** 3
        assert A.nnz == 7 * n - 6 * s * s
        e = np.ones(s)
        T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
        I = sp.identity(s)
        K = sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)
        assert abs(A - K).max() == 0


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_reference_factor_1d_closed_form():
    # 1D Dirichlet chain: every eliminated vertex has two neighbours, the
    # single sampled edge is the exact Schur complement, so G is the exact
    # Cholesky factor: G_ii = sqrt((i+2)/(i+1)), G_{i+1,i} = -sqrt((i+1)/(i+2)).
    n = 20
    cp, ri, v = O.ref_rchol_seq(lap1d(n), 123)
    G = sp.csc_matrix((v, ri, cp), shape=(n, n)).toarray()
    i = np.arange(n)
    np.testing.assert_allclose(np.diag(G), np.sqrt((i + 2) / (i + 1)), rtol=1e-14)
    np.testing.assert_allclose(np.diag(G, -1), -np.sqrt((i[:-1] + 1) / (i[:-1] + 2)), rtol=1e-14)
    np.testing.assert_allclose(G @ G.T, lap1d(n).toarray(), atol=1e-13)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_reference_factor_invariants_and_seed():
    A = O.ref_laplace_3d(6)
    n = A.shape[0]
    g1 = O.ref_rchol_seq(A, 5)
    g2 = O.ref_rchol_seq(A, 5)
    g3 = O.ref_rchol_seq(A, 6)
    assert all(np.array_equal(a, b) for a, b in zip(g1, g2))
    assert not np.array_equal(g1[2], g3[2])
    cp, ri, v = g1
    G = sp.csc_matrix((v, ri, cp), shape=(n, n))
    for j in range(n):
        rows = ri[cp[j]:cp[j + 1]]
        vals = v[cp[j]:cp[j + 1]]
        assert rows[0] == j and vals[0] > 0           # diagonal first, positive
        assert np.all(np.diff(rows) > 0)               # ascending rows
        assert np.all(vals[1:] < 0)                    # off-diagonals negative
        assert vals.sum() >= -1e-12                     # column-sum invariant
    assert sp.triu(G, 1).nnz == 0


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_reference_reorder_matches_scipy_permutation():
    A = O.ref_laplace_3d(4)
    p = np.random.default_rng(0).permutation(A.shape[0])
    B = O.ref_reorder(A, p)
    assert abs(B - A[p][:, p]).max() == 0
    assert all(np.all(np.diff(B.indices[B.indptr[i]:B.indptr[i + 1]]) > 0)
               for i in range(B.shape[0]))


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_golden(name):
    A, Gs, d = load(name)
    np.testing.assert_array_equal(O.precond(Gs, d["b"]), d["z"])
    x, relres, itr = O.pcg(A, d["b"], float(d["tol"]), 5000, Gs)
    assert itr == int(d["itr"])
    np.testing.assert_allclose(x, d["x"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", FIXTURES)
def test_golden_factor_is_sane(name):
    A, Gs, d = load(name)
    n = A.shape[0]
    G = sp.csc_matrix((Gs[2], Gs[1], Gs[0]), shape=(n, n))
    assert sp.triu(G, 1).nnz == 0
    # G G^T approximates A: the preconditioned spectrum is bounded
    r = d["b"]
    z = O.precond(Gs, r)
    assert np.dot(r, z) > 0


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("name", FIXTURES)
def test_golden_factor_regenerates_from_reference(name):
    A, Gs, d = load(name)
    perm, blocks, seed = d["perm"], d["blocks"], int(d["seed"])
    n = A.shape[0]
    if blocks.size == 1:
        g = O.ref_rchol_seq(A, seed)
    else:
        # fixtures hold the permuted A; undo to feed the parallel path
        inv = np.empty(n, dtype=np.int64)
        inv[perm.astype(np.int64)] = np.arange(n)
        A0 = A[inv][:, inv]
        g = O.ref_rchol_nd(A0, perm, blocks, seed)
    for a, b in zip(g, Gs):
        np.testing.assert_array_equal(a, b)


def test_prepared_oracle_equals_the_plain_calls():
    """oracle.Prepared (arrays converted once, for the bench's CPU timing)
    runs the same C code as oracle.pcg / oracle.precond: identical bits on
    one thread (OpenMP's reduction order varies between calls)."""
    A = O.ref_laplace_3d(6) if O.ref_available() else None
    if A is None:
        pytest.skip("reference build absent")
    Gs = O.ref_rchol_seq(A, 4)
    b = np.random.default_rng(9).uniform(-1, 1, A.shape[0])
    P = O.Prepared(A, Gs)
    O.set_threads(1)
    try:
        x1, r1, i1 = P.pcg(b, 1e-10, 100)
        x2, r2, i2 = O.pcg(A, b, 1e-10, 100, Gs)
    finally:
        O.set_threads(os.cpu_count() or 1)
    assert i1 == i2 and r1 == r2
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(P.precond(b), O.precond(Gs, b))
