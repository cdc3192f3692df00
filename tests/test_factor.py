"""Host factorization, orderings and generators (librchol_factor.so) -- CPU.

Pins: bit-identity with the reference's own factorization on the golden
fixtures (and live against oracle/_ref where it is built), the MIS order's
factor against the reference's elimination on the order MIS chose,
unbiasedness of
the sampled factor (E[G G^T] = A), the 1D closed form, and structural
invariants of both orderings.
"""
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

from rchol_amd import factor as F
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz")))


def load(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    n = d["a_indptr"].size - 1
    A = sp.csr_matrix((d["a_data"], d["a_indices"].astype(np.int64),
                       d["a_indptr"].astype(np.int64)), shape=(n, n))
    return A, d


def check_factor_structure(Gs, n):
    cp, ri, v = Gs
    assert cp[0] == 0 and cp[-1] == ri.size and np.all(np.diff(cp.astype(np.int64)) >= 1)
    for j in range(n):
        rows = ri[cp[j]:cp[j + 1]]
        vals = v[cp[j]:cp[j + 1]]
        assert rows[0] == j and vals[0] > 0
        assert np.all(np.diff(rows.astype(np.int64)) > 0)
        assert np.all(vals[1:] < 0)
        assert vals.sum() >= -1e-12 * vals[0]  # column-sum invariant


@pytest.mark.parametrize("name", FIXTURES)
def test_static_factor_is_bit_identical_to_reference_fixture(name):
    Ap, d = load(name)
    n = Ap.shape[0]
    perm = d["perm"].astype(np.int64)
    blocks = d["blocks"]
    if blocks.size == 1:
        Gs, po, _ = F.factor(Ap, None, [n], "static", int(d["seed"]))
    else:
        inv = np.empty(n, dtype=np.int64)
        inv[perm] = np.arange(n)
        Gs, po, _ = F.factor(Ap[inv][:, inv], perm, blocks, "static", int(d["seed"]))
        assert np.array_equal(po, perm)
    for a, b in zip(Gs, (d["g_colptr"], d["g_rowidx"], d["g_val"])):
        np.testing.assert_array_equal(a, b)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("case", ["seq16", "nd8_var20", "ndgraph16_lap14"])
def test_static_factor_matches_reference_live(case):
    if case == "seq16":
        A = F.laplace_3d(16)
        g_ref = O.ref_rchol_seq(A, 4242)
        g, _, _ = F.factor(A, None, [A.shape[0]], "static", 4242)
    elif case == "nd8_var20":
        A = F.varcoef_3d(20, 18, 16, 1.5, 1e-3, 9)
        perm, blocks = F.nd_grid(20, 18, 16, 4)
        g_ref = O.ref_rchol_nd(A, perm, blocks, 77)
        g, _, _ = F.factor(A, perm, blocks, "static", 77)
    else:
        A = F.laplace_3d(14)
        perm, blocks = F.nd_graph(A, 5)
        g_ref = O.ref_rchol_nd(A, perm, blocks, 1)
        g, _, _ = F.factor(A, perm, blocks, "static", 1)
    for a, b in zip(g, g_ref):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("case", ["var12", "var16", "var2d64"])
def test_mis_factor_is_the_reference_algorithm_on_its_own_order(case):
    """The MIS order picks the elimination order inside each block while it
    eliminates; the factor it returns is then the reference's elimination
    (rchol_lap, built from its sources when present, else this package's
    bit-identical static path) run on that same order and seed: the same
    sampled pattern, values equal up to the order in which duplicate fill is
    summed (1e-12 relative).  Variable coefficients: with the constant
    Laplacian's tied weights that summation order can reorder a tie and so
    the sampling."""
    if case == "var12":
        A, grid, depth = F.varcoef_3d(12, 12, 12, 1.5, 1e-3, 3), (12, 12, 12), 3
    elif case == "var16":
        A, grid, depth = F.varcoef_3d(16, 16, 16, 1.5, 1e-3, 8), (16, 16, 16), 3
    else:
        A, grid, depth = F.varcoef_3d(64, 64, 1, 1.0, 0.0, 5), (64, 64, 1), 4
    n = A.shape[0]
    perm, blocks = F.nd_grid(*grid, depth)
    Gm, po, _ = F.factor(A, perm, blocks, "mis", 7)
    if O.ref_available():
        Gr = O.ref_rchol_nd(A, po, blocks, 7)
    else:
        Gr, _, _ = F.factor(A, po, blocks, "static", 7)

    def mat(G):
        cp, ri, gv = G
        return sp.csc_matrix((gv, ri.astype(np.int64), cp.astype(np.int64)), shape=(n, n))
    Mm, Mr = mat(Gm), mat(Gr)
    assert Mm.nnz == Mr.nnz
    assert (Mm != 0).multiply(Mr != 0).sum() == Mm.nnz          # same pattern
    assert abs(Mm - Mr).max() <= 1e-12 * abs(Mr).max()


def test_laplace_3d_matches_reference_closed_form():
    for s in (2, 4, 7):
        A = F.laplace_3d(s)
        n = s ** 3
        assert A.nnz == 7 * n - 6 * s * s
        e = np.ones(s)
        T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
        I = sp.identity(s)
        K = sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)
        assert abs(A - K).max() == 0


def test_one_dimensional_chain_gives_exact_cholesky():
    n = 25
    e = np.ones(n)
    A = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], format="csr")
    for order in ("static", "mis"):
        Gs, po, _ = F.factor(A, None, [n], order, 3)
        G = F.lower(Gs).toarray()
        P = po.astype(np.int64)
        np.testing.assert_allclose(G @ G.T, A.toarray()[np.ix_(P, P)], atol=1e-12)


@pytest.mark.parametrize("order", ["static", "mis"])
def test_factor_is_unbiased(order):
    # E[G G^T] = A: each sampled clique has the exact Schur complement as
    # its mean (weight w_i (W - C_i)/W on an edge drawn with probability
    # w_j / (W - C_i)).  A wrong weight or draw rule biases the mean.
    A = F.varcoef_3d(3, 3, 3, 1.0, 0.1, 2)
    n = A.shape[0]
    perm, blocks = F.nd_grid(3, 3, 3, 2)
    S = 600
    acc = np.zeros((n, n))
    for s in range(S):
        Gs, po, _ = F.factor(A, perm, blocks, order, s + 1)
        G = F.lower(Gs).toarray()
        P = po.astype(np.int64)
        M = np.empty((n, n))
        M[np.ix_(P, P)] = G @ G.T  # back to the original numbering
        acc += M
    err = np.linalg.norm(acc / S - A.toarray()) / np.linalg.norm(A.toarray())
    assert err < 0.01, err


@pytest.mark.parametrize("order", ["static", "mis"])
def test_factor_structure_and_ordering(order):
    A = F.varcoef_3d(10, 9, 8, 1.5, 1e-3, 4)
    n = A.shape[0]
    perm, blocks = F.nd_grid(10, 9, 8, 3)
    Gs, po, _ = F.factor(A, perm, blocks, order, 12)
    assert np.array_equal(np.sort(po.astype(np.int64)), np.arange(n))
    check_factor_structure(Gs, n)
    # ND blocks are kept: the set of vertices in every block is unchanged
    cuts = np.concatenate([[0], np.cumsum(blocks.astype(np.int64))])
    for a, b in zip(cuts[:-1], cuts[1:]):
        assert set(po[a:b].tolist()) == set(perm[a:b].tolist())
    # determinism
    Gs2, po2, _ = F.factor(A, perm, blocks, order, 12)
    assert np.array_equal(po, po2) and all(np.array_equal(x, y) for x, y in zip(Gs, Gs2))


def sweep_levels(Gs):
    cp, ri, _ = Gs
    n = cp.size - 1
    lev = np.zeros(n, dtype=np.int64)
    for j in range(n):
        rows = ri[cp[j] + 1:cp[j + 1]].astype(np.int64)
        if rows.size:
            np.maximum.at(lev, rows, lev[j] + 1)
    return int(lev.max()) + 1


def test_mis_order_shortens_the_sweeps():
    A = F.varcoef_3d(24, 24, 24, 1.5, 1e-3, 1)
    perm, blocks = F.nd_grid(24, 24, 24, 4)
    g_static, _, _ = F.factor(A, perm, blocks, "static", 1)
    g_mis, _, _ = F.factor(A, perm, blocks, "mis", 1)
    ls, lm = sweep_levels(g_static), sweep_levels(g_mis)
    assert lm * 8 < ls, (ls, lm)
    assert g_mis[1].size < g_static[1].size


def test_nd_grid_contract():
    perm, blocks = F.nd_grid(8, 6, 5, 3)
    n = 8 * 6 * 5
    assert np.array_equal(np.sort(perm.astype(np.int64)), np.arange(n))
    assert blocks.size == 7 and blocks.sum() == n
    # separators separate: no edge between the two leaves of one parent
    A = F.laplace_3d(1)  # noqa: F841  (library loaded)
    cuts = np.concatenate([[0], np.cumsum(blocks.astype(np.int64))])
    leaf0 = set(perm[cuts[0]:cuts[1]].tolist())
    leaf1 = set(perm[cuts[1]:cuts[2]].tolist())
    for v in leaf0:
        x, y, z = v % 8, (v // 8) % 6, v // 48
        for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            X, Y, Z = x + dx, y + dy, z + dz
            if 0 <= X < 8 and 0 <= Y < 6 and 0 <= Z < 5:
                assert X + 8 * Y + 48 * Z not in leaf1


def test_nd_graph_contract_and_errors():
    A = F.varcoef_3d(9, 9, 9, 1.0, 0.0, 3)
    n = A.shape[0]
    perm, blocks = F.nd_graph(A, 4)
    assert np.array_equal(np.sort(perm.astype(np.int64)), np.arange(n))
    assert blocks.size == 15 and blocks.sum() == n
    Ap = A[perm.astype(np.int64)][:, perm.astype(np.int64)]
    cuts = np.concatenate([[0], np.cumsum(blocks.astype(np.int64))])
    # leaves of different subtrees are not coupled
    B = Ap[cuts[0]:cuts[1], cuts[1]:cuts[2]]
    assert B.nnz == 0
    with pytest.raises(Exception, match="too many threads"):
        F.nd_graph(sp.identity(3, format="csr"), 4)


def test_rchol_python_interface():
    A = F.laplace_3d(10)
    n = A.shape[0]
    G, perm, part = F.rchol(A)
    assert np.array_equal(perm, np.arange(n)) and list(part) == [0, n + 1]
    assert sp.triu(G, 1).nnz == 0
    G2, perm2, part2 = F.rchol(A, 4, grid=(10, 10, 10))
    assert part2[-1] == n + 1 and part2.size == 8
    # reuse of an ordering (ex_reuse_partition.py:37)
    B = A + 1e-3 * sp.identity(n)
    G3, perm3, part3 = F.rchol(B, 4, perm2, part2)
    assert np.array_equal(perm3, perm2)
    with pytest.raises(Exception):
        F.rchol(A, 0)
    with pytest.raises(Exception):
        F.rchol(A, 2, perm=np.arange(n))
