"""The drop-in surfaces above the C ABI, on the GPU.

- Python: ex_laplace.py's call sequence -- G, p, _ = rchol(A, ...)
  (python/rchol/rchol.py:8-45), then x = pcg(A[p][:, p], b[p], tol, maxit,
  G, G.T) (python/rchol/util.py:32-61) and x_orig[p] = x
  (python/rchol/ex_laplace.py:14-31) -- checked against the oracle on the
  same inputs and against the unpermuted system.
- C++: the pcg(A, b, tol, maxit, G, x, relres, itr) constructor of
  c++/util/pcg.hpp:15-16 (include/rchol_amd_pcg.hpp), through the example
  driver that mirrors c++/ex_laplace.cpp, against the oracle on the same
  factor and b (x and the iteration count); and the reference's own
  c++/ex_laplace.cpp, compiled unchanged against a forwarding pcg.hpp,
  against the oracle on its own seeded factor and right-hand side.
- The device-pointer entry points (torch tensors) against their host twins:
  the same kernels run, so the results are bit-identical; a device batch
  that is not 4, 8, 16 or 32 columns wide is refused.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import rchol_amd
from rchol_amd import factor as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if rchol_amd.device_count() < 1:
        pytest.skip("no HIP device")


def test_python_dropin_follows_ex_laplace():
    n = 12
    A = F.laplace_3d(n)
    b = np.random.default_rng(3).uniform(-1.0, 1.0, A.shape[0])
    G, perm, part = F.rchol(A, 4, grid=(n, n, n))
    assert part[-1] == A.shape[0] + 1
    p = perm.astype(np.int64)
    Ap = A[p][:, p].tocsr()
    print("case python_dropin", flush=True)
    x, relres, niters = rchol_amd.pcg(Ap, b[p], 1e-10, 500, G, G.T.tocsr())
    xo, ro, io = O.pcg(Ap, b[p], 1e-10, 500, rchol_amd.stored_from_lower(G))
    assert abs(niters - io) <= 1
    assert rel(x, xo) <= 1e-9
    x_orig = np.empty_like(x)
    x_orig[p] = x
    true = np.linalg.norm(A @ x_orig - b) / np.linalg.norm(b)
    assert relres == pytest.approx(true, rel=1e-6, abs=1e-14)
    assert true <= 2e-10


def test_cpp_pcg_shim_example(tmp_path):
    """The pcg(...) constructor of include/rchol_amd_pcg.hpp through the
    example driver, against the oracle on the same factor and the same b."""
    exe = os.path.join(ROOT, "rchol_amd", "csrc", "bin", "ex_laplace")
    assert os.path.exists(exe), "the C++ example is built by make -C rchol_amd/csrc all"
    side, leaves, seed, tol = 16, 4, 1, 1e-10
    print("case cpp_shim vs oracle", flush=True)
    n = side ** 3
    b = np.random.default_rng(5).uniform(-1.0, 1.0, n)
    bfile, xfile = tmp_path / "b.bin", tmp_path / "x.bin"
    b.tofile(bfile)
    out = subprocess.run([exe, "-n", str(side), "-t", str(leaves), "-tol", str(tol),
                          "-maxit", "500", "-seed", str(seed), "-bfile", str(bfile),
                          "-xfile", str(xfile)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    fill = float(re.search(r"Fill-in ratio: (\S+)", out.stdout).group(1))
    itr = int(re.search(r"# CG iterations: (\d+)", out.stdout).group(1))
    relres = float(re.search(r"Relative residual: (\S+)", out.stdout).group(1))
    x = np.fromfile(xfile, dtype=np.float64)
    # the same factor on the host: static order, ND with 4 leaves, same seed
    A = F.laplace_3d(side)
    perm, blocks = F.nd_grid(side, side, side, 3)
    Gs, po, _ = F.factor(A, perm, blocks, "static", seed)
    P = po.astype(np.int64)
    Ap = F.permute(A, po)
    assert fill == pytest.approx(2.0 * Gs[0][-1] / Ap.nnz, rel=2e-3)
    xo_p, ro, io = O.pcg(Ap, b[P], tol, 500, Gs)
    xo = np.empty(n)
    xo[P] = xo_p
    print(f"cpp shim: itr {itr} oracle {io}, x {rel(x, xo):.1e}", flush=True)
    assert abs(itr - io) <= 1
    assert rel(x, xo) <= 1e-8
    assert 0.0 < relres <= 2 * tol
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 2 * tol


def test_reference_driver_unchanged_on_the_dropin():
    """The reference's own c++/ex_laplace.cpp, compiled unchanged with the
    forwarding pcg.hpp (oracle/Makefile target dropin), solves on this
    engine: its factor (rchol, seeded through the driver's random_device)
    and right-hand side (util.hpp:50-54 rand) are rebuilt on the host and
    the oracle runs the same system."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from std_random import reference_rand
    exe = os.path.join(ROOT, "oracle", "_ref", "ex_laplace_dropin")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/ex_laplace_dropin not built (needs the reference sources)")
    if not O.ref_available():
        pytest.skip("oracle/_ref/librchol_ref.so not built")
    for side, seed in ((12, 3), (20, 11)):
        print(f"case reference driver n={side} seed={seed}", flush=True)
        env = dict(os.environ, RCHOL_DROPIN_SEED=str(seed))
        out = subprocess.run([exe, "-n", str(side)], capture_output=True, text=True,
                             timeout=120, env=env)
        assert out.returncode == 0, out.stderr
        fill = float(re.search(r"Fill-in ratio: (\S+)", out.stdout).group(1))
        itr = int(re.search(r"# CG iterations: (\d+)", out.stdout).group(1))
        relres = float(re.search(r"Relative residual: (\S+)", out.stdout).group(1))
        A = O.ref_laplace_3d(side)
        N = A.shape[0]
        Gs = O.ref_rchol_seq(A, seed)
        b = reference_rand(N, seed)
        assert fill == pytest.approx(2.0 * Gs[0][-1] / A.nnz, rel=2e-3)
        xo, ro, io = O.pcg(A, b, 1e-6, 200, Gs)   # ex_laplace.cpp:37-38
        print(f"reference driver: itr {itr} oracle {io}, relres {relres:.3g} oracle {ro:.3g}",
              flush=True)
        assert abs(itr - io) <= 1
        assert relres <= 1e-6
        if itr == io:
            assert relres == pytest.approx(ro, rel=1e-2)


def test_device_entry_points_match_host_ones():
    import torch

    d = np.load(os.path.join(GOLD, "lap3d_8_nd4.npz"))
    n = d["a_indptr"].size - 1
    A = sp.csr_matrix((d["a_data"], d["a_indices"], d["a_indptr"]), shape=(n, n))
    Gs = (d["g_colptr"], d["g_rowidx"], d["g_val"])
    b = d["b"]
    print("case device_entry_points", flush=True)
    for prec in (rchol_amd.FP64, rchol_amd.FP32):
        with rchol_amd.Solver(A, Gs, prec) as s:
            xh, rh, ih = s.pcg(b, 1e-10, 500)
            bt = torch.from_numpy(b).cuda()
            xt = torch.zeros_like(bt)
            torch.cuda.synchronize()
            rd, idv = s.pcg_device(bt.data_ptr(), xt.data_ptr(), 1e-10, 500)
            assert idv == ih and rd == rh
            assert np.array_equal(xt.cpu().numpy(), xh)
            zt = torch.empty_like(bt)
            s.precond_device(bt.data_ptr(), zt.data_ptr())
            assert np.array_equal(zt.cpu().numpy(), s.precond(b))
            # device batches take 4, 8, 16 or 32 interleaved columns
            B = np.ascontiguousarray(np.stack([b, 2.0 * b, -b[::-1], b * b], axis=1))
            Xh, rrh, ith = s.pcg_batch(B, 1e-10, 500)
            Bt = torch.from_numpy(B).cuda()
            Xt = torch.zeros_like(Bt)
            torch.cuda.synchronize()
            rrd, itd = s.pcg_batch_device(4, Bt.data_ptr(), Xt.data_ptr(), 1e-10, 500)
            assert np.array_equal(itd, ith) and np.array_equal(rrd, rrh)
            assert np.array_equal(Xt.cpu().numpy(), Xh)
            with pytest.raises(RuntimeError, match="4, 8, 16 or 32 columns"):
                s.pcg_batch_device(3, Bt.data_ptr(), Xt.data_ptr(), 1e-10, 500)


def test_reuse_partition_sequence_updates_values_in_place():
    # python/ex_reuse_partition.py: A, then B = A + 1e-3 I (same sparsity)
    # factored with the first ordering; here the second solve keeps the
    # device residency of A's pattern and only replaces its values
    n = 20
    A = F.laplace_3d(n)
    N = A.shape[0]
    b = np.random.default_rng(11).random(N)
    G, perm, part = F.rchol(A, 2)
    p = perm.astype(np.int64)
    B = (A + 1e-3 * sp.identity(N)).tocsr()
    L = F.rchol(B, 2, perm, part)[0]
    Aperm, Bperm = A[p][:, p].tocsr(), B[p][:, p].tocsr()
    print("case reuse_partition", flush=True)
    with rchol_amd.Solver(Aperm, G) as s:
        x1, r1, i1 = s.pcg(b[p], 1e-6, 200)
        assert r1 <= 1e-6
        s.update_A_values(Bperm)
        s.upload_G(L)
        x2, r2, i2 = s.pcg(b[p], 1e-6, 200)
        with pytest.raises(ValueError, match="pattern"):
            s.update_A_values(sp.identity(N, format="csr"))
        # the C ABI itself refuses a value count that is not the resident nnz
        short = np.ones(Bperm.nnz - 1)
        assert s._lib.rchol_amd_update_A_values(
            s._h, short.size, short.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == -1
        assert b"differs from the resident" in s._lib.rchol_amd_last_error()
    with rchol_amd.Solver(Bperm, L) as fresh:
        x3, r3, i3 = fresh.pcg(b[p], 1e-6, 200)
    assert i2 == i3 and r2 == r3 and np.array_equal(x2, x3)
    xo, ro, io = O.pcg(Bperm, b[p], 1e-6, 200, rchol_amd.stored_from_lower(L))
    assert abs(i2 - io) <= 1 and rel(x2, xo) <= 1e-9


def test_set_perm_solves_in_the_original_numbering():
    """rchol_amd_set_perm: with A(P,P) and its factor resident, pcg, precond
    and the batched solve take and return vectors in the original numbering
    (b(P) gathered, x(P) = x' scattered on the device), against the oracle
    on the permuted system; clearing P restores the permuted numbering."""
    print("case set_perm 20^3 ND 8 leaves", flush=True)
    A = F.varcoef_3d(20, 20, 20, 1.5, 1e-3, 12)
    n = A.shape[0]
    perm, blocks = F.nd_grid(20, 20, 20, 4)
    Gs, po, _ = F.factor(A, perm, blocks, "mis", 3)
    P = po.astype(np.int64)
    Ap = F.permute(A, po)
    b = np.random.default_rng(21).uniform(-1, 1, n)
    r = np.random.default_rng(22).uniform(-1, 1, n)
    xo_p, _, io = O.pcg(Ap, b[P], 1e-10, 500, Gs)
    zo = np.empty(n)
    zo[P] = O.precond(Gs, r[P])
    with rchol_amd.Solver(Ap, Gs, rchol_amd.FP64) as s:
        s.set_perm(po)
        x, relres, itr = s.pcg(b, 1e-10, 500)
        z = s.precond(r)
        B = np.stack([b, r, b - r, 2 * r], axis=1)
        X, rr, it = s.pcg_batch(B, 1e-10, 500)
        s.set_perm(None)
        xp, _, _ = s.pcg(b[P], 1e-10, 500)
    x_orig = np.empty(n)
    x_orig[P] = xo_p
    assert abs(itr - io) <= 1
    assert rel(x, x_orig) <= 1e-8
    assert rel(z, zo) <= 1e-12
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-9
    for c in range(B.shape[1]):
        assert np.linalg.norm(A @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c]) <= 1e-9
    np.testing.assert_allclose(X[:, 0], x, rtol=0, atol=1e-9 * np.abs(x).max())
    assert rel(xp, xo_p) <= 1e-8
