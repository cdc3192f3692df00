"""GPU parity against the oracle at the benchmark sizes.

The oracle is the C restatement of c++/util/pcg.cpp:57-159 (oracle/pcg_oracle.c),
fed the same A(perm, perm), stored factor and b as the engine.  Bars:
  * FP64 preconditioner application z = G^-T G^-1 r within 1e-12 relative
    (2-norm) of the oracle's;
  * PCG iteration count to a relative residual of 1e-10 within +-1 of the
    oracle's, and the solution within 1e-8 relative of the oracle's;
  * FP32 factor: application within 1e-5 relative, solves FP64-verified to
    1e-9 (true residual on the host).
Workloads:
  C1   32^3 Laplacian, sequential rchol, natural order
  C2   128^3 Laplacian, nested dissection with 64 leaves, both the reference's
       order inside the blocks ("static") and independent-set layers ("mis")
  C3   256^3 variable-coefficient Poisson, 128 leaves, mis
  C4a  the C3 system with 32 right-hand sides
  C4b  48 independent 2D 5-point systems, sides 64..1024, one launch per stage
"""
import numpy as np
import pytest

import rchol_amd
from rchol_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PRECOND_RTOL = 1e-12
FP32_PRECOND_RTOL = 1e-5
X_RTOL = 1e-8
TOL = 1e-10


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def true_relres(A, x, b):
    return float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if rchol_amd.device_count() < 1:
        pytest.skip("no HIP device")
    O.set_threads(16)


def check_solve(A, Gs, b, tag, maxit=5000):
    """Full FP64 solve on both sides: itr +-1, x, reported relres = true relres."""
    xo, ro, io = O.pcg(A, b, TOL, maxit, Gs)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP64) as s:
        z = s.precond(b)
        x, relres, itr = s.pcg(b, TOL, maxit)
        st = s.stats()
    zo = O.precond(Gs, b)
    print(f"{tag}: itr {itr} oracle {io}; precond {rel(z, zo):.1e}; x {rel(x, xo):.1e}; "
          f"chunked {st['chunked']} ({st['chunks']} chunks), {st['last_solve_ms']:.1f} ms",
          flush=True)
    assert rel(z, zo) <= PRECOND_RTOL
    assert abs(itr - io) <= 1
    assert rel(x, xo) <= X_RTOL
    assert relres == pytest.approx(true_relres(A, x, b), rel=1e-6)
    assert relres <= TOL * 10


def test_c1_laplace_32_natural_order():
    print("case C1 32^3 natural order", flush=True)
    A, Gs, po = W.laplace_natural(32)
    np.testing.assert_array_equal(po, np.arange(A.shape[0]))
    check_solve(A, Gs, W.rhs(A.shape[0]), "C1")


@pytest.mark.parametrize("order", ["static", "mis"])
def test_c2_laplace_128_nd64(order):
    print(f"case C2 128^3 ND 64 order={order}", flush=True)
    A, Gs, po = W.laplace_nd(128, 64, order)
    b = W.rhs(A.shape[0])[po.astype(np.int64)]
    check_solve(A, Gs, b, f"C2/{order}")
    # FP32 factor on the same system
    zo = O.precond(Gs, b)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP32) as s:
        z32 = s.precond(b)
        x, relres, itr = s.pcg(b, TOL, 5000)
    assert rel(z32, zo) <= FP32_PRECOND_RTOL
    assert true_relres(A, x, b) <= TOL * 10


@pytest.fixture(scope="module")
def c3():
    print("building C3: 256^3 variable-coefficient Poisson, ND 128 leaves, mis", flush=True)
    A, Gs, po = W.varcoef_nd(256, 128, "mis")
    b = W.rhs(A.shape[0])[po.astype(np.int64)]
    return A, Gs, b


def test_c3_precond_fp64_and_fp32(c3):
    A, Gs, b = c3
    print("case C3 precond", flush=True)
    zo = O.precond(Gs, b)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP64) as s:
        z = s.precond(b)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP32) as s:
        z32 = s.precond(b)
    print(f"C3 precond fp64 {rel(z, zo):.1e} fp32 {rel(z32, zo):.1e}", flush=True)
    assert rel(z, zo) <= PRECOND_RTOL
    assert rel(z32, zo) <= FP32_PRECOND_RTOL


def test_c3_pcg_iterations_match_oracle(c3):
    A, Gs, b = c3
    print("case C3 pcg", flush=True)
    xo, ro, io = O.pcg(A, b, TOL, 2000, Gs)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP64) as s:
        x, relres, itr = s.pcg(b, TOL, 2000)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP32) as s:
        x32, relres32, itr32 = s.pcg(b, TOL, 2000)
        # the fenced reduction handshake gives the same bits at this size too
        s.set_option(rchol_amd.solver.OPT_STRICT_SYNC, 1)
        x32s, relres32s, itr32s = s.pcg(b, TOL, 2000)
    print(f"C3 itr fp64 {itr} mixed {itr32} oracle {io}; x {rel(x, xo):.1e}", flush=True)
    assert abs(itr - io) <= 1
    assert rel(x, xo) <= X_RTOL
    assert true_relres(A, x, b) <= TOL * 10
    assert true_relres(A, x32, b) <= TOL * 10
    np.testing.assert_array_equal(x32s, x32)
    assert (relres32s, itr32s) == (relres32, itr32)


def test_c4a_batch_of_32_against_oracle(c3):
    A, Gs, b = c3
    print("case C4a 256^3 x 32", flush=True)
    n, k = A.shape[0], 32
    B = np.random.default_rng(11).uniform(-1.0, 1.0, (n, k))
    with rchol_amd.Solver(A, Gs, rchol_amd.FP64) as s:
        Z = s.precond_batch(B)
        X, rr, it = s.pcg_batch(B, TOL, 2000)
        for c in range(k):
            assert rel(Z[:, c], s.precond(B[:, c].copy())) <= PRECOND_RTOL
    for c in (0, 17, 31):
        assert rel(Z[:, c], O.precond(Gs, B[:, c].copy())) <= PRECOND_RTOL
    for c in (0, 31):
        xo, ro, io = O.pcg(A, B[:, c].copy(), TOL, 2000, Gs)
        print(f"C4a column {c}: itr {it[c]} oracle {io}, x {rel(X[:, c], xo):.1e}", flush=True)
        assert abs(int(it[c]) - io) <= 1
        assert rel(X[:, c], xo) <= X_RTOL
    for c in range(k):
        assert true_relres(A, X[:, c], B[:, c]) <= TOL * 10
        assert rr[c] == pytest.approx(true_relres(A, X[:, c], B[:, c]), rel=1e-6)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP32) as s:
        Z32 = s.precond_batch(B)
        X32, rr32, it32 = s.pcg_batch(B, TOL, 2000)
    for c in (0, 17, 31):
        assert rel(Z32[:, c], Z[:, c]) <= FP32_PRECOND_RTOL
    for c in range(k):
        assert true_relres(A, X32[:, c], B[:, c]) <= TOL * 10


def test_c4b_uneven_batch_of_48_against_oracle():
    print("case C4b 48 uneven 2D systems", flush=True)
    systems, bs, sides = W.uneven_2d(48)
    Au, Gu, off = rchol_amd.stack_systems(systems)
    b = np.concatenate(bs)
    print(f"C4b n_total={int(off[-1])} sides {int(sides.min())}..{int(sides.max())}", flush=True)
    with rchol_amd.Solver(Au, Gu, rchol_amd.FP64) as s:
        s.set_segments(off)
        x, rr, it = s.pcg_segments(b, TOL, 5000)
    with rchol_amd.Solver(Au, Gu, rchol_amd.FP32) as s:
        s.set_segments(off)
        x32, rr32, it32 = s.pcg_segments(b, TOL, 5000)
    worst = 0
    for k, (A, Gs) in enumerate(systems):
        lo, hi = int(off[k]), int(off[k + 1])
        xo, ro, io = O.pcg(A, bs[k], TOL, 5000, Gs)
        worst = max(worst, abs(int(it[k]) - io))
        assert abs(int(it[k]) - io) <= 1, (k, int(sides[k]), int(it[k]), io)
        assert rel(x[lo:hi], xo) <= X_RTOL, k
        assert true_relres(A, x32[lo:hi], bs[k]) <= TOL * 10, k
        assert rr[k] == pytest.approx(true_relres(A, x[lo:hi], bs[k]), rel=1e-6)
    print(f"C4b: max |itr - oracle| = {worst}", flush=True)
