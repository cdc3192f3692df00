"""SDD -> SDDM adapter (rchol_amd.sdd), after matlab/rchol/sdd_to_sddm.m and
matlab/ex_sdd.m."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from rchol_amd import sdd as S


def test_sdd_3d_is_sdd_with_positive_couplings():
    A = S.sdd_3d(5)
    off = A - sp.diags(A.diagonal())
    assert off.max() > 0 and off.min() < 0
    assert np.all(A.diagonal() >= abs(off).sum(axis=1).A1 - 1e-12)
    assert abs(A - A.T).max() == 0


def test_extension_is_sddm_and_recovers_the_solution():
    A = S.sdd_3d(4)
    n = A.shape[0]
    Ae = S.sdd_to_sddm(A)
    off = Ae - sp.diags(Ae.diagonal())
    assert off.max() <= 0                                   # M-matrix sign pattern
    assert np.all(Ae.diagonal() >= abs(off).sum(axis=1).A1 - 1e-12)
    b = np.random.default_rng(0).uniform(size=n)
    xe = spl.spsolve(sp.csc_matrix(Ae), S.extend_rhs(b))
    x = S.recover(xe)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)


@pytest.mark.gpu
def test_solve_sdd_on_gpu():
    import rchol_amd
    if rchol_amd.device_count() < 1:
        pytest.skip("no HIP device")
    print("case solve_sdd", flush=True)
    A = S.sdd_3d(10)
    b = np.random.default_rng(1).uniform(size=A.shape[0])
    x, relres, itr = S.solve_sdd(A, b, tol=1e-10)
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)
    assert relres <= 1e-9 and itr > 0
