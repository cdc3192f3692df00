"""GPU checks at a size where every stage runs on a full grid (64^3
variable-coefficient Poisson, 262k rows): the reduction handshake, the task
cut and the oracle at a size the C oracle still finishes in seconds."""
import numpy as np
import pytest

import rchol_amd
from rchol_amd import factor as F
from rchol_amd.solver import OPT_STRICT_SYNC, OPT_TASK_WAVES
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    if rchol_amd.device_count() < 1:
        pytest.skip("no HIP device")
    side = 64
    A = F.varcoef_3d(side, side, side, 1.5, 1e-3, 11)
    perm, blocks = F.nd_grid(side, side, side, 4)
    Gs, po, _ = F.factor(A, perm, blocks, "mis", 1)
    Ap = F.permute(A, po)
    b = np.random.default_rng(5).uniform(-1.0, 1.0, A.shape[0])
    return Ap, Gs, b


@pytest.mark.parametrize("precision", [rchol_amd.FP64, rchol_amd.FP32])
def test_reduction_handshakes_agree_bitwise(system, precision):
    """The default reduction handshake (value-returning atomics) and the
    fenced one must produce the same bits: every dot product is summed in
    block order either way, so any partial read before it was published
    would show up as a different x."""
    Ap, Gs, b = system
    print(f"case handshake prec={precision}", flush=True)
    out = {}
    with rchol_amd.Solver(Ap, Gs, precision) as s:
        for strict in (1, 0, 1, 0):
            s.set_option(OPT_STRICT_SYNC, strict)
            x, relres, itr = s.pcg(b, 1e-10, 500)
            if strict in out:
                np.testing.assert_array_equal(x, out[strict][0])
            out[strict] = (x, relres, itr)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2]


def test_precond_matches_oracle_at_scale(system):
    Ap, Gs, b = system
    print("case precond oracle 64^3", flush=True)
    z_ref = O.precond(Gs, b)
    with rchol_amd.Solver(Ap, Gs) as s:
        z = s.precond(b)
    assert np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref) <= 1e-12


@pytest.mark.parametrize("waves", [64, 512])
def test_stream_cut_at_scale(system, waves):
    """A cut planned for few waves makes most wide-level rows stream tasks;
    the application agrees with the default cut to rounding."""
    Ap, Gs, b = system
    print(f"case stream cut waves={waves}", flush=True)
    with rchol_amd.Solver(Ap, Gs) as s:
        z0 = s.precond(b)
        x0, r0, i0 = s.pcg(b, 1e-10, 500)
        s.set_option(OPT_TASK_WAVES, waves)
        z1 = s.precond(b)
        x1, r1, i1 = s.pcg(b, 1e-10, 500)
    assert np.linalg.norm(z1 - z0) / np.linalg.norm(z0) <= 1e-12
    assert abs(i1 - i0) <= 1
    true = np.linalg.norm(Ap @ x1 - b) / np.linalg.norm(b)
    assert true <= 1e-9


@pytest.mark.parametrize("k", [3, 8, 32])
def test_batch_columns_match_single_solves(system, k):
    """The K-column sweeps (their own task cut and wave lists) against the
    single-right-hand-side path, column by column, at a size where every
    level is cut for the full grid."""
    Ap, Gs, b = system
    print(f"case batch-vs-single k={k}", flush=True)
    rng = np.random.default_rng(100 + k)
    B = rng.uniform(-1.0, 1.0, (Ap.shape[0], k))
    with rchol_amd.Solver(Ap, Gs) as s:
        Z = s.precond_batch(B)
        X, rr, it = s.pcg_batch(B, 1e-10, 500)
        for c in range(k):
            z1 = s.precond(B[:, c].copy())
            assert np.linalg.norm(Z[:, c] - z1) / np.linalg.norm(z1) <= 1e-12
            if c < 3:
                x1, r1, i1 = s.pcg(B[:, c].copy(), 1e-10, 500)
                assert abs(int(it[c]) - i1) <= 1
                assert np.linalg.norm(X[:, c] - x1) / np.linalg.norm(x1) <= 1e-8
            true = np.linalg.norm(Ap @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c])
            assert true <= 1e-9
