"""The error path of the bounded waits, forced from the host: with the wait
bound lowered to one tick (RCHOL_AMD_OPT_WAIT_TICKS), a solve whose sweeps
must wait reports RCHOL_AMD_ESTALL instead of hanging; every kernel still
exits, and once the bound is restored the same context solves correctly
(checked against the oracle).  Both sweep forms: level by level and in
chunks.  The factor is a 1D chain in natural order, so every row waits for
the one before it; in the band case every row also reads the row two
before it, so the level-by-level waits go through a proxy (an entry two
levels down) before every lane polls its own pending values."""
import numpy as np
import pytest
import scipy.sparse as sp

import rchol_amd
from rchol_amd import factor as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if rchol_amd.device_count() < 1:
        pytest.skip("no HIP device")


def chain(n, seed, band=1):
    rng = np.random.default_rng(seed)
    A = sp.csr_matrix((n, n))
    d = np.zeros(n) + 1e-3
    for k in range(1, band + 1):
        w = rng.uniform(0.5, 2.0, n - k)
        A = A + sp.diags([-w, -w], [-k, k], shape=(n, n))
        d[:-k] += w
        d[k:] += w
    A = (A + sp.diags(d)).tocsr()
    A.sort_indices()
    Gs, _, _ = F.factor(A, None, [n], "static", seed)
    return A, Gs


@pytest.mark.parametrize("chunked,n,band", [(0, 20000, 1), (1, 70000, 1), (0, 20000, 2)])
def test_lowered_wait_bound_reports_a_stall_then_recovers(chunked, n, band):
    print(f"case stall chunked={chunked} n={n} band={band}", flush=True)
    A, Gs = chain(n, 3 + chunked, band)
    b = np.random.default_rng(9).uniform(-1, 1, n)
    xo, _, io = O.pcg(A, b, 1e-10, 50, Gs)
    with rchol_amd.Solver(A, Gs, rchol_amd.FP64, chunked=chunked) as s:
        assert s.stats()["chunked"] == chunked
        s.set_option(rchol_amd.solver.OPT_WAIT_TICKS, 1)
        try:
            with pytest.raises(RuntimeError, match=r"error -6: triangular sweep stalled"):
                s.pcg(b, 1e-10, 50)
        finally:
            s.set_option(rchol_amd.solver.OPT_WAIT_TICKS, 0)
        x, relres, itr = s.pcg(b, 1e-10, 50)
    assert abs(itr - io) <= 1
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    assert relres <= 1e-9
