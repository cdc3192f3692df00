"""The allocation-free preconditioner entry point
(rchol_amd_precond_device_noalloc) against rchol_amd_precond_device, through
the C ABI: the same sweep kernels run on the engine's static scratch instead
of the context's work vectors, so the results are bit-identical; both are
also held to the oracle's 1e-12.  Two calls in a row, the second on a
smaller system (the scratch keeps the first call's contents), with and
without a permutation, FP64 and FP32 factors; a factor beyond the scratch
is refused with a clear error.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import rchol_amd
from rchol_amd import factor as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def system(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    n = d["a_indptr"].size - 1
    A = sp.csr_matrix((d["a_data"], d["a_indices"], d["a_indptr"]), shape=(n, n))
    return A, (d["g_colptr"], d["g_rowidx"], d["g_val"])


@pytest.fixture(scope="module")
def torch():
    if rchol_amd.device_count() < 1:
        pytest.skip("no HIP device")
    import torch as T
    assert T.cuda.is_available(), "torch sees no GPU next to the engine"
    return T


@pytest.mark.parametrize("use_perm", [False, True])
@pytest.mark.parametrize("precision", [rchol_amd.FP64, rchol_amd.FP32])
def test_noalloc_precond_matches_precond_device(torch, precision, use_perm):
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(3 + precision + 2 * use_perm)
    # larger system first, then a smaller one on the same GPU's scratch
    for name in ("lap3d_12_nd8", "lap3d_8_nd4", "lap3d_12_nd8"):
        A, Gs = system(name)
        n = A.shape[0]
        print(f"case noalloc {name} n={n} prec={precision} perm={use_perm}", flush=True)
        perm = rng.permutation(n) if use_perm else None
        with rchol_amd.Solver(A, Gs, precision) as s:
            if use_perm:
                s.set_perm(perm)
            r = rng.uniform(-1.0, 1.0, n)
            rt = torch.tensor(r, device=dev)
            z1 = torch.full_like(rt, np.nan)
            z2 = torch.full_like(rt, np.nan)
            torch.cuda.synchronize()
            s.precond_device(rt.data_ptr(), z1.data_ptr())
            s.precond_device_noalloc(rt.data_ptr(), z2.data_ptr())
            s.precond_device_noalloc(rt.data_ptr(), z2.data_ptr())
            a, b = z1.cpu().numpy(), z2.cpu().numpy()
        np.testing.assert_array_equal(b, a)
        # the oracle on the factor's numbering: z = P^T M^-1 P r
        if use_perm:
            zo = np.empty(n)
            zo[perm] = O.precond(Gs, r[perm])
        else:
            zo = O.precond(Gs, r)
        assert rel(b, zo) <= (1e-12 if precision == rchol_amd.FP64 else 1e-5)


def test_noalloc_refuses_a_factor_beyond_its_scratch(torch):
    # a diagonal factor of 3 M rows with a permutation needs 24 bytes per row
    n = 3_000_000
    print(f"case noalloc refuses n={n}", flush=True)
    cp = np.arange(n + 1, dtype=np.uint64)
    ri = np.arange(n, dtype=np.uint64)
    gv = np.full(n, 2.0)
    dev = torch.device("cuda", 0)
    with rchol_amd.Solver(None, (cp, ri, gv), rchol_amd.FP64) as s:
        s.set_perm(np.arange(n)[::-1].copy())
        r = torch.ones(n, dtype=torch.float64, device=dev)
        z = torch.empty_like(r)
        torch.cuda.synchronize()
        with pytest.raises(RuntimeError, match="scratch"):
            s.precond_device_noalloc(r.data_ptr(), z.data_ptr())
        # the allocating entry point still serves it
        s.precond_device(r.data_ptr(), z.data_ptr())
        assert torch.all(z == 0.25).item()
