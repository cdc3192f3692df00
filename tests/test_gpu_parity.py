"""GPU parity of the HIP path against the oracle (CPU restatement of
c++/util/pcg.cpp), through the C ABI.

Bars: FP64 preconditioner application within 1e-12
relative (2-norm) of the oracle; PCG iteration count to 1e-10 within +-1 of
the oracle; SpMV within 1e-13 relative.  The FP32 factor path is held to
1e-5 relative on the application and to an FP64-verified 1e-10 solve.
"""
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

import rchol_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz")))

PRECOND_RTOL = 1e-12
SPMV_RTOL = 1e-13


def load(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    n = d["a_indptr"].size - 1
    A = sp.csr_matrix((d["a_data"], d["a_indices"], d["a_indptr"]), shape=(n, n))
    return A, (d["g_colptr"], d["g_rowidx"], d["g_val"]), d


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if rchol_amd.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.mark.parametrize("name", FIXTURES)
def test_precond_golden(name):
    A, Gs, d = load(name)
    print(f"case precond_golden {name} n={A.shape[0]}", flush=True)
    with rchol_amd.Solver(A, Gs) as s:
        z = s.precond(d["b"])
        assert rel(z, d["z"]) <= PRECOND_RTOL
        # repeated applications are identical (deterministic sweeps)
        z2 = s.precond(d["b"])
        np.testing.assert_array_equal(z, z2)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("lanes", [0, 8])
def test_spmv_golden(name, lanes):
    A, Gs, d = load(name)
    print(f"case spmv_golden {name} lanes={lanes}", flush=True)
    with rchol_amd.Solver(A, Gs, spmv_lanes=lanes) as s:
        y = s.spmv(d["b"])
        assert rel(y, d["Ab"]) <= SPMV_RTOL


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("graph_iters", [0, 16])
@pytest.mark.parametrize("lanes", [0, 8])
def test_pcg_golden(name, graph_iters, lanes):
    A, Gs, d = load(name)
    print(f"case pcg_golden {name} graph_iters={graph_iters} spmv_lanes={lanes}", flush=True)
    with rchol_amd.Solver(A, Gs, graph_iters=graph_iters, spmv_lanes=lanes) as s:
        x, relres, itr = s.pcg(d["b"], float(d["tol"]), 5000)
    assert abs(itr - int(d["itr"])) <= 1
    assert relres <= float(d["tol"]) * 10
    assert rel(x, d["x"]) <= 1e-8
    true = np.linalg.norm(A @ x - d["b"]) / np.linalg.norm(d["b"])
    assert relres == pytest.approx(true, rel=1e-6)


@pytest.mark.parametrize("name", FIXTURES)
def test_precond_fp32_golden(name):
    A, Gs, d = load(name)
    print(f"case precond_fp32 {name}", flush=True)
    with rchol_amd.Solver(A, Gs, precision=rchol_amd.FP32) as s:
        z = s.precond(d["b"])
    assert rel(z, d["z"]) <= 1e-5


@pytest.mark.parametrize("name", FIXTURES)
def test_pcg_fp32_factor_reaches_fp64_tolerance(name):
    A, Gs, d = load(name)
    print(f"case pcg_fp32 {name}", flush=True)
    with rchol_amd.Solver(A, Gs, precision=rchol_amd.FP32) as s:
        x, relres, itr = s.pcg(d["b"], 1e-10, 5000)
    true = np.linalg.norm(A @ x - d["b"]) / np.linalg.norm(d["b"])
    assert true <= 1e-9
    assert itr <= int(d["itr"]) + 5


def test_pcg_reference_semantics_edges():
    A, Gs, d = load("lap3d_4_seq")
    b = d["b"]
    print("case pcg_edges", flush=True)
    with rchol_amd.Solver(A, Gs) as s:
        # maxit cut-off is silent (pcg.cpp:82)
        x, relres, itr = s.pcg(b, 1e-14, 3)
        xo, ro, io = O.pcg(A, b, 1e-14, 3, Gs)
        assert itr == io == 3
        assert rel(x, xo) <= 1e-11
        # maxit = 0: no iteration, x stays zero, relres = 1
        x, relres, itr = s.pcg(b, 1e-10, 0)
        assert itr == 0 and np.all(x == 0) and relres == pytest.approx(1.0)
        # pre-sized x is an accumulator while r starts at b (pcg.cpp:67-71)
        x0 = np.full(b.size, 0.25)
        x1, _, i1 = s.pcg(b, 1e-10, 7, x0=x0)
        xo1, _, _ = O.pcg(A, b, 1e-10, 7, Gs, x0=x0)
        assert rel(x1, xo1) <= 1e-10
        # repeated solves on the same context reuse the residency
        for _ in range(3):
            x, relres, itr = s.pcg(b, 1e-10, 500)
            assert abs(itr - int(d["itr"])) <= 1


def test_exact_factor_one_iteration():
    n = 64
    e = np.ones(n)
    A = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], format="csr")
    Gc = sp.csc_matrix(np.linalg.cholesky(A.toarray()))
    b = np.random.default_rng(0).uniform(-1, 1, n)
    print("case exact_factor", flush=True)
    with rchol_amd.Solver(A, (Gc.indptr, Gc.indices, Gc.data)) as s:
        x, relres, itr = s.pcg(b, 1e-10, 10)
    assert itr == 1
    assert rel(x, np.linalg.solve(A.toarray(), b)) <= 1e-12


def test_diagonal_factor_and_single_row():
    print("case diagonal_factor", flush=True)
    A = sp.csr_matrix(np.array([[4.0]]))
    with rchol_amd.Solver(A, (np.array([0, 1]), np.array([0]), np.array([2.0]))) as s:
        x, relres, itr = s.pcg(np.array([8.0]), 1e-10, 10)
        assert itr == 1 and x[0] == pytest.approx(2.0)
        z = s.precond(np.array([8.0]))
        assert z[0] == pytest.approx(2.0)


def test_upload_rejects_bad_factor():
    A, Gs, d = load("lap3d_4_seq")
    print("case bad_factor", flush=True)
    with rchol_amd.Solver(A) as s:
        cp, ri, v = (np.array(a, copy=True) for a in Gs)
        v[cp[3]] = 0.0  # zero pivot of column 3
        with pytest.raises(RuntimeError, match="error -5"):
            s.upload_G((cp, ri, v))
        cp, ri, v = (np.array(a, copy=True) for a in Gs)
        ri[cp[5]], ri[cp[5] + 1] = ri[cp[5] + 1], ri[cp[5]]  # diagonal not first is fine,
        s.upload_G((cp, ri, v))                               # but must still be lower
        bad = (np.array([0, 1, 2]), np.array([1, 1]), np.array([1.0, 1.0]))
        with pytest.raises(RuntimeError):
            s.upload_G(bad)  # column 0 misses its diagonal


@pytest.mark.parametrize("name", ["lap3d_8_nd4", "varcoef_6_nd2", "lap3d_12_nd8"])
@pytest.mark.parametrize("k", [3, 8, 32])
@pytest.mark.parametrize("precision", [rchol_amd.FP64, rchol_amd.FP32])
def test_batch_pcg_matches_single_column_oracle(name, k, precision):
    A, Gs, d = load(name)
    n = A.shape[0]
    print(f"case batch {name} k={k} prec={precision}", flush=True)
    rng = np.random.default_rng(k)
    B = rng.uniform(-1, 1, (n, k))
    B[:, 0] = d["b"]
    with rchol_amd.Solver(A, Gs, precision=precision) as s:
        X, rr, it = s.pcg_batch(B, 1e-10, 5000)
        Z = s.precond_batch(B)
    for c in range(k):
        xo, ro, io = O.pcg(A, B[:, c], 1e-10, 5000, Gs)
        zo = O.precond(Gs, B[:, c])
        true = np.linalg.norm(A @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c])
        if precision == rchol_amd.FP64:
            assert abs(int(it[c]) - io) <= 1
            assert rel(Z[:, c], zo) <= PRECOND_RTOL
            assert rel(X[:, c], xo) <= 1e-8
        else:
            assert rel(Z[:, c], zo) <= 1e-5
            assert int(it[c]) <= io + 5
        assert true <= 1e-9
        assert rr[c] == pytest.approx(true, rel=1e-6)


def test_batch_columns_stop_independently():
    A, Gs, d = load("lap3d_8_nd4")
    n = A.shape[0]
    print("case batch_masks", flush=True)
    B = np.zeros((n, 4))
    B[:, 0] = d["b"]
    B[:, 1] = 1e-3 * d["b"]  # same direction, same iteration count
    B[:, 3] = np.random.default_rng(0).uniform(-1, 1, n)
    # column 2 is zero: no iteration, x stays 0
    with rchol_amd.Solver(A, Gs) as s:
        X, rr, it = s.pcg_batch(B, 1e-10, 5000)
        x0, _, i0 = s.pcg(d["b"], 1e-10, 5000)
    assert it[2] == 0 and np.all(X[:, 2] == 0)
    assert abs(int(it[0]) - i0) <= 1 and abs(int(it[1]) - i0) <= 1
    assert rel(X[:, 0], x0) <= 1e-9


@pytest.mark.parametrize("period", [5, 50])
def test_fp32_factor_with_residual_replacement(period):
    A, Gs, d = load("lap3d_12_nd8")
    print(f"case resid_replace period={period}", flush=True)
    with rchol_amd.Solver(A, Gs, precision=rchol_amd.FP32) as s:
        s.set_option(rchol_amd.solver.OPT_RESID_REPLACE, period)
        x, relres, itr = s.pcg(d["b"], 1e-10, 5000)
    true = np.linalg.norm(A @ x - d["b"]) / np.linalg.norm(d["b"])
    assert true <= 1e-9 and relres == pytest.approx(true, rel=1e-6)
    assert itr <= int(d["itr"]) + 5


@pytest.mark.parametrize("precision", [rchol_amd.FP64, rchol_amd.FP32])
@pytest.mark.parametrize("lanes", [0, 8])
@pytest.mark.parametrize("strict", [0, 1])
def test_uneven_batch_matches_per_system_oracle(precision, lanes, strict):
    from rchol_amd import factor as F
    print(f"case segments prec={precision} spmv_lanes={lanes} strict={strict}", flush=True)
    rng = np.random.default_rng(48)
    systems = []
    for s, side in enumerate([6, 13, 9, 20, 7]):
        A = F.varcoef_3d(side, side + 1, 1, 1.0, 0.0, 100 + s)
        p, bl = F.nd_grid(side, side + 1, 1, 2)
        Gs, po, _ = F.factor(A, p, bl, "mis", s + 1)
        P = po.astype(np.int64)
        systems.append((A[P][:, P].tocsr(), Gs))
    Au, Gu, off = rchol_amd.stack_systems(systems)
    b = rng.uniform(-1, 1, Au.shape[0])
    with rchol_amd.Solver(Au, Gu, precision=precision, spmv_lanes=lanes) as s:
        s.set_option(rchol_amd.solver.OPT_STRICT_SYNC, strict)
        s.set_segments(off)
        x, rr, it = s.pcg_segments(b, 1e-10, 2000)
        x2, rr2, it2 = s.pcg_segments(b, 1e-10, 2000)  # graph replay, same result
    np.testing.assert_array_equal(x, x2)
    for k, (A, Gs) in enumerate(systems):
        lo, hi = int(off[k]), int(off[k + 1])
        xo, ro, io = O.pcg(A, b[lo:hi], 1e-10, 2000, Gs)
        true = np.linalg.norm(A @ x[lo:hi] - b[lo:hi]) / np.linalg.norm(b[lo:hi])
        assert true <= 1e-9 and rr[k] == pytest.approx(true, rel=1e-6)
        if precision == rchol_amd.FP64:
            assert abs(int(it[k]) - io) <= 1
            assert rel(x[lo:hi], xo) <= 1e-8
        else:
            assert int(it[k]) <= io + 5


def test_autotune_keeps_results():
    A, Gs, d = load("lap3d_12_nd8")
    print("case autotune", flush=True)
    with rchol_amd.Solver(A, Gs) as s:
        t = s.autotune(reps=2)
        assert t["trsv_lanes"] in (1, 2, 4, 8) and t["spmv_lanes"] in (0, 4, 8, 16)
        z = s.precond(d["b"])
        x, relres, itr = s.pcg(d["b"], 1e-10, 5000)
    assert rel(z, d["z"]) <= PRECOND_RTOL
    assert abs(itr - int(d["itr"])) <= 1


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("width", [16, 1 << 20])
@pytest.mark.parametrize("precision", [rchol_amd.FP64, rchol_amd.FP32])
def test_split_sweeps_match_oracle(name, width, precision):
    A, Gs, d = load(name)
    print(f"case split {name} width={width} prec={precision}", flush=True)
    with rchol_amd.Solver(A, Gs, precision=precision) as s:
        s.set_option(rchol_amd.solver.OPT_SERIAL_WIDTH, width)
        z = s.precond(d["b"])
        x, relres, itr = s.pcg(d["b"], 1e-10, 5000)
    true = np.linalg.norm(A @ x - d["b"]) / np.linalg.norm(d["b"])
    if precision == rchol_amd.FP64:
        assert rel(z, d["z"]) <= PRECOND_RTOL
        assert abs(itr - int(d["itr"])) <= 1
    else:
        assert rel(z, d["z"]) <= 1e-5
        assert itr <= int(d["itr"]) + 5
    assert true <= 1e-9 and relres == pytest.approx(true, rel=1e-6)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("waves", [1, 3])
@pytest.mark.parametrize("precision", [rchol_amd.FP64, rchol_amd.FP32])
def test_stream_tasks_match_oracle(name, waves, precision):
    """A task cut planned for very few waves turns the short rows of every
    level into stream tasks (coalesced entry block, LDS row sums) and the
    rest into multi-pass lane tasks."""
    A, Gs, d = load(name)
    print(f"case stream {name} waves={waves} prec={precision}", flush=True)
    with rchol_amd.Solver(A, Gs, precision=precision) as s:
        s.set_option(rchol_amd.solver.OPT_TASK_WAVES, waves)
        z = s.precond(d["b"])
        x, relres, itr = s.pcg(d["b"], 1e-10, 5000)
        st = s.stats()
    assert 0 < st["fwd_tasks"] <= A.shape[0] and 0 < st["bwd_tasks"] <= A.shape[0]
    true = np.linalg.norm(A @ x - d["b"]) / np.linalg.norm(d["b"])
    if precision == rchol_amd.FP64:
        assert rel(z, d["z"]) <= PRECOND_RTOL
        assert abs(itr - int(d["itr"])) <= 1
    else:
        assert rel(z, d["z"]) <= 1e-5
        assert itr <= int(d["itr"]) + 5
    assert true <= 1e-9 and relres == pytest.approx(true, rel=1e-6)


def test_option_values_are_validated():
    A, Gs, d = load("lap3d_8_nd4")
    print("case options", flush=True)
    lib = rchol_amd._lib.engine()
    with rchol_amd.Solver(A, Gs) as s:
        h = s._h
        for key, bad in ((rchol_amd.solver.OPT_TRSV_LANES, 3), (rchol_amd.solver.OPT_TRSV_LANES, 128),
                         (rchol_amd.solver.OPT_SPMV_LANES, 5), (rchol_amd.solver.OPT_TASK_WAVES, -1),
                         (rchol_amd.solver.OPT_SERIAL_WIDTH, -1), (99, 0)):
            assert lib.rchol_amd_set_option(h, key, float(bad)) == -1
        for key, good in ((rchol_amd.solver.OPT_TRSV_LANES, 64), (rchol_amd.solver.OPT_SPMV_LANES, 0),
                          (rchol_amd.solver.OPT_TASK_WAVES, 0), (rchol_amd.solver.OPT_STRICT_SYNC, 1)):
            assert lib.rchol_amd_set_option(h, key, float(good)) == 0
        z = s.precond(d["b"])
        x, relres, itr = s.pcg(d["b"], 1e-10, 5000)
    assert rel(z, d["z"]) <= PRECOND_RTOL
    assert abs(itr - int(d["itr"])) <= 1


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("precision", [rchol_amd.FP64, rchol_amd.FP32])
def test_chunked_sweeps_match_oracle(name, precision):
    """The chunked sweeps (forced on these small factors) against the
    oracle: application, repeatability, PCG iterations and solution."""
    A, Gs, d = load(name)
    print(f"case chunked {name} prec={precision}", flush=True)
    with rchol_amd.Solver(A, Gs, precision, chunked=1) as s:
        assert s.stats()["chunked"] == 1
        z = s.precond(d["b"])
        np.testing.assert_array_equal(z, s.precond(d["b"]))
        x, relres, itr = s.pcg(d["b"], float(d["tol"]), 5000)
    if precision == rchol_amd.FP64:
        assert rel(z, d["z"]) <= PRECOND_RTOL
        assert abs(itr - int(d["itr"])) <= 1
        assert rel(x, d["x"]) <= 1e-8
    else:
        assert rel(z, d["z"]) <= 1e-5
    true = np.linalg.norm(A @ x - d["b"]) / np.linalg.norm(d["b"])
    assert true <= float(d["tol"]) * 10
    assert relres == pytest.approx(true, rel=1e-6)
