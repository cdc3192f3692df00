"""Differentiable solve (rchol_amd.autograd): gradient checks.

CPU: the adjoint formula dL/dw_e = -(lam_u - lam_v)(x_u - x_v), dL/db = lam
against central finite differences of a dense FP64 solve (pins the math the
edge-gradient kernel implements, and the grid Laplacian assembly).
GPU: the HIP forward/adjoint against FP64 central finite differences on a
16^3 instance (20 seeded directions, 1e-6 relative), and the mixed-precision
gradient against an FP64 adjoint.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

from rchol_amd import factor as F


def softplus(t):
    return np.log1p(np.exp(-np.abs(t))) + np.maximum(t, 0)


def refined_solve(A, b, steps=4):
    """FP64 LU solve refined with extended-precision (long double) residuals,
    so that finite differences of the loss are not swamped by the solve's
    rounding (cond(A) * eps)."""
    import scipy.sparse as sp
    As = sp.csr_matrix(A)
    lu = spl.splu(sp.csc_matrix(As))
    data = As.data.astype(np.longdouble)
    idx, ptr = As.indices, As.indptr
    bl = np.asarray(b, dtype=np.longdouble)
    x = lu.solve(np.asarray(b, dtype=np.float64)).astype(np.longdouble)
    for _ in range(steps):
        Ax = np.add.reduceat(data * x[idx], ptr[:-1])
        r = bl - Ax
        x = x + lu.solve(r.astype(np.float64)).astype(np.longdouble)
    return x


def dense_laplacian(dims, w, shift):
    u, v = F.grid_edges(*dims)
    n = dims[0] * dims[1] * dims[2]
    A = np.diag(np.full(n, shift))
    for e in range(u.size):
        a, b = u[e], v[e]
        A[a, a] += w[e]
        A[b, b] += w[e]
        A[a, b] -= w[e]
        A[b, a] -= w[e]
    return A, u, v


def test_grid_laplacian_assembly_matches_dense():
    dims = (4, 3, 2)
    rng = np.random.default_rng(0)
    u, v = F.grid_edges(*dims)
    assert u.size == 3 * 24 - 12 - 6 - 8
    w = rng.uniform(0.5, 2.0, u.size)
    Ad, _, _ = dense_laplacian(dims, w, 0.3)
    A = F.grid_laplacian(*dims, w, 0.3)
    np.testing.assert_allclose(A.toarray(), Ad, rtol=0, atol=1e-15)


def test_adjoint_formula_matches_finite_differences():
    dims = (3, 3, 2)
    rng = np.random.default_rng(1)
    u, v = F.grid_edges(*dims)
    theta = rng.standard_normal(u.size)
    n = 18
    b = rng.uniform(-1, 1, n)
    xs = rng.uniform(-1, 1, n)

    def loss(th, bb):
        A, _, _ = dense_laplacian(dims, softplus(th), 1e-2)
        x = refined_solve(A, bb)
        return np.sum((x - xs) ** 2), np.asarray(x, dtype=np.float64), A

    L0, x, A = loss(theta, b)
    lam = np.linalg.solve(A, 2 * (x - xs))
    gw = -(lam[u] - lam[v]) * (x[u] - x[v])
    gtheta = gw / (1 + np.exp(-theta))  # softplus'
    h = 1e-5
    for k in range(5):
        d = rng.standard_normal(u.size)
        fd = float((loss(theta + h * d, b)[0] - loss(theta - h * d, b)[0]) / (2 * h))
        assert abs(fd - gtheta @ d) <= 1e-6 * abs(fd)
        db = rng.standard_normal(n)
        fd = float((loss(theta, b + h * db)[0] - loss(theta, b - h * db)[0]) / (2 * h))
        assert abs(fd - lam @ db) <= 1e-6 * abs(fd)


@pytest.mark.gpu
def test_gpu_gradient_matches_fp64_finite_differences():
    torch = pytest.importorskip("torch")
    import rchol_amd
    from rchol_amd import autograd as AG
    if rchol_amd.device_count() < 1 or not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("case autograd_fd 16^3", flush=True)
    dims = (16, 16, 16)
    n = 4096
    rng = np.random.default_rng(2026)
    prob = AG.GridLaplacian(*dims, shift=1e-2, leaves=8, precision=rchol_amd.FP32, tol=1e-12)
    theta0 = rng.standard_normal(prob.m)
    b0 = rng.uniform(-1, 1, n)
    xs = rng.uniform(-1, 1, n)
    dev = torch.device("cuda", 0)
    theta = torch.tensor(theta0, device=dev, requires_grad=True)
    b = torch.tensor(b0, device=dev, requires_grad=True)
    xstar = torch.tensor(xs, device=dev)
    w = torch.nn.functional.softplus(theta)
    x = AG.solve(w, b, prob)
    loss = ((x - xstar) ** 2).sum()
    loss.backward()
    gth = theta.grad.cpu().numpy()
    gb = b.grad.cpu().numpy()

    def ref_loss(th, bb):
        A = F.grid_laplacian(*dims, softplus(th), 1e-2)
        xx = refined_solve(A, bb)
        return np.sum((xx - xs) ** 2)

    assert abs(loss.item() - float(ref_loss(theta0, b0))) <= 1e-8 * float(ref_loss(theta0, b0))
    h = 1e-5
    drng = np.random.default_rng(20)
    for k in range(20):
        d = drng.standard_normal(prob.m)
        fd = float((ref_loss(theta0 + h * d, b0) - ref_loss(theta0 - h * d, b0)) / (2 * h))
        assert abs(gth @ d - fd) <= 1e-6 * abs(fd), (k, gth @ d, fd)
    for k in range(3):
        d = drng.standard_normal(n)
        fd = float((ref_loss(theta0, b0 + h * d) - ref_loss(theta0, b0 - h * d)) / (2 * h))
        assert abs(gb @ d - fd) <= 1e-6 * abs(fd)


@pytest.mark.gpu
def test_gpu_mixed_gradient_matches_fp64_adjoint():
    torch = pytest.importorskip("torch")
    import rchol_amd
    from rchol_amd import autograd as AG
    if rchol_amd.device_count() < 1 or not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("case autograd_adjoint 32^3", flush=True)
    dims = (32, 32, 32)
    n = 32 ** 3
    rng = np.random.default_rng(5)
    dev = torch.device("cuda", 0)
    theta0 = rng.standard_normal(3 * n - 3 * 32 * 32)
    b0 = rng.uniform(-1, 1, n)
    xs = rng.uniform(-1, 1, n)
    grads = {}
    for prec in (rchol_amd.FP32, rchol_amd.FP64):
        prob = AG.GridLaplacian(*dims, shift=1e-2, leaves=8, precision=prec, tol=1e-12)
        theta = torch.tensor(theta0, device=dev, requires_grad=True)
        x = AG.solve(torch.nn.functional.softplus(theta), torch.tensor(b0, device=dev), prob)
        ((x - torch.tensor(xs, device=dev)) ** 2).sum().backward()
        grads[prec] = theta.grad.cpu().numpy()
    g32, g64 = grads[rchol_amd.FP32], grads[rchol_amd.FP64]
    assert np.linalg.norm(g32 - g64) <= 1e-6 * np.linalg.norm(g64)
