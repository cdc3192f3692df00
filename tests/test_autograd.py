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


def softplus_ld(t):
    t = np.asarray(t, dtype=np.longdouble)
    return np.log1p(np.exp(-np.abs(t))) + np.maximum(t, 0)


def laplacian_matvec_ld(u, v, w, shift, x):
    """(L(w) + shift I) x in long double, from the edge list."""
    d = x[u] - x[v]
    f = w * d
    y = shift * x
    np.add.at(y, u, f)
    np.add.at(y, v, -f)
    return y


def refined_solve(A, matvec_ld, b, steps=6):
    """Solve with an FP64 LU of A refined against the long-double operator
    matvec_ld (weights and residuals in extended precision), so that central
    differences of the loss are not swamped by FP64 assembly and solve
    rounding (~cond(A) * eps)."""
    import scipy.sparse as sp
    lu = spl.splu(sp.csc_matrix(A))
    bl = np.asarray(b, dtype=np.longdouble)
    x = lu.solve(np.asarray(b, dtype=np.float64)).astype(np.longdouble)
    for _ in range(steps):
        r = bl - matvec_ld(x)
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


@pytest.mark.parametrize("dims", [(4, 3, 2), (7, 6, 5), (5, 1, 3), (1, 4, 1), (1, 1, 1)])
def test_grid_laplacian_assembly_matches_dense(dims):
    rng = np.random.default_rng(0)
    u, v = F.grid_edges(*dims)
    nx, ny, nz = dims
    assert u.size == 3 * nx * ny * nz - nx * ny - ny * nz - nx * nz
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
        w = softplus_ld(th)
        x = refined_solve(A, lambda y: laplacian_matvec_ld(u, v, w, np.longdouble(1e-2), y), bb)
        return np.sum((x - xs) ** 2), np.asarray(x, dtype=np.float64), A

    L0, x, A = loss(theta, b)
    lam = np.linalg.solve(A, 2 * (x - xs))
    gw = -(lam[u] - lam[v]) * (x[u] - x[v])
    gtheta = gw / (1 + np.exp(-theta))  # softplus'
    h = 1e-5
    for k in range(5):
        d = rng.standard_normal(u.size)
        d = d.astype(np.longdouble)
        fd = float((loss(theta + h * d, b)[0] - loss(theta - h * d, b)[0]) / (2 * h))
        assert abs(fd - gtheta @ d.astype(np.float64)) <= 1e-6 * abs(fd)
        db = rng.standard_normal(n)
        db = db.astype(np.longdouble)
        fd = float((loss(theta, b + h * db)[0] - loss(theta, b - h * db)[0]) / (2 * h))
        assert abs(fd - lam @ db.astype(np.float64)) <= 1e-6 * abs(fd)


@pytest.mark.gpu
def test_gpu_gradient_matches_fp64_finite_differences():
    torch = pytest.importorskip("torch")
    import rchol_amd
    from rchol_amd import autograd as AG
    if rchol_amd.device_count() < 1:
        pytest.skip("no GPU")
    # the engine and torch must share one HIP runtime in the process
    assert torch.cuda.is_available(), "torch sees no GPU next to the engine"
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

    ue, ve = F.grid_edges(*dims)

    def ref_loss(th, bb):
        A = F.grid_laplacian(*dims, softplus(th), 1e-2)
        w = softplus_ld(th)
        xx = refined_solve(A, lambda y: laplacian_matvec_ld(ue, ve, w, np.longdouble(1e-2), y), bb)
        return np.sum((xx - xs) ** 2)

    assert abs(loss.item() - float(ref_loss(theta0, b0))) <= 1e-8 * float(ref_loss(theta0, b0))
    h = 1e-5
    drng = np.random.default_rng(20)
    for k in range(20):
        d = drng.standard_normal(prob.m)
        dl = d.astype(np.longdouble)
        fd = float((ref_loss(theta0 + h * dl, b0) - ref_loss(theta0 - h * dl, b0)) / (2 * h))
        assert abs(gth @ d - fd) <= 1e-6 * abs(fd), (k, gth @ d, fd)
    for k in range(3):
        d = drng.standard_normal(n)
        dl = d.astype(np.longdouble)
        fd = float((ref_loss(theta0, b0 + h * dl) - ref_loss(theta0, b0 - h * dl)) / (2 * h))
        assert abs(gb @ d - fd) <= 1e-6 * abs(fd)


@pytest.mark.gpu
def test_gpu_mixed_gradient_matches_fp64_adjoint():
    torch = pytest.importorskip("torch")
    import rchol_amd
    from rchol_amd import autograd as AG
    if rchol_amd.device_count() < 1:
        pytest.skip("no GPU")
    # the engine and torch must share one HIP runtime in the process
    assert torch.cuda.is_available(), "torch sees no GPU next to the engine"
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


@pytest.mark.gpu
def test_gpu_two_pending_solves_keep_their_own_factors():
    """Two forward solves with different weights before one backward: each
    adjoint must run with the factor of its own A(w).  Checked against the
    dense FP64 adjoint gradient of both solves."""
    torch = pytest.importorskip("torch")
    import rchol_amd
    from rchol_amd import autograd as AG
    if rchol_amd.device_count() < 1:
        pytest.skip("no GPU")
    assert torch.cuda.is_available(), "torch sees no GPU next to the engine"
    print("case autograd two pending solves 8^3", flush=True)
    dims = (8, 8, 8)
    n = 512
    rng = np.random.default_rng(77)
    prob = AG.GridLaplacian(*dims, shift=1e-2, leaves=4, precision=rchol_amd.FP64, tol=1e-13)
    dev = torch.device("cuda", 0)
    th1 = rng.standard_normal(prob.m)
    th2 = rng.standard_normal(prob.m) + 1.0
    b0 = rng.uniform(-1, 1, n)
    s1, s2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    t1 = torch.tensor(th1, device=dev, requires_grad=True)
    t2 = torch.tensor(th2, device=dev, requires_grad=True)
    b = torch.tensor(b0, device=dev)
    x1 = AG.solve(torch.nn.functional.softplus(t1), b, prob)
    x2 = AG.solve(torch.nn.functional.softplus(t2), b, prob)
    loss = ((x1 - torch.tensor(s1, device=dev)) ** 2).sum() + \
        ((x2 - torch.tensor(s2, device=dev)) ** 2).sum()
    loss.backward()
    u, v = F.grid_edges(*dims)
    for th, tt, xs in ((th1, t1, s1), (th2, t2, s2)):
        A, _, _ = dense_laplacian(dims, softplus(th), 1e-2)
        x = np.linalg.solve(A, b0)
        lam = np.linalg.solve(A, 2 * (x - xs))
        g_ref = -(lam[u] - lam[v]) * (x[u] - x[v]) / (1 + np.exp(-th))
        g = tt.grad.cpu().numpy()
        assert np.linalg.norm(g - g_ref) <= 1e-8 * np.linalg.norm(g_ref)


@pytest.mark.gpu
def test_gpu_mixed_gradient_matches_fp64_adjoint_at_full_size():
    """The benchmark's 100^3 instance: the mixed-precision forward + adjoint
    gradient against the FP64 adjoint (PCG to 1e-13), within 1e-6."""
    torch = pytest.importorskip("torch")
    import rchol_amd
    from rchol_amd import autograd as AG
    if rchol_amd.device_count() < 1:
        pytest.skip("no GPU")
    assert torch.cuda.is_available(), "torch sees no GPU next to the engine"
    print("case autograd adjoint 100^3", flush=True)
    side = 100
    n = side ** 3
    dev = torch.device("cuda", 0)
    b0 = np.random.default_rng(5).uniform(-1, 1, n)
    xs = np.random.default_rng(2).uniform(-1, 1, n)
    grads = {}
    for prec, tol in ((rchol_amd.FP32, 1e-10), (rchol_amd.FP64, 1e-13)):
        prob = AG.GridLaplacian(side, side, side, shift=1e-2, leaves=16, precision=prec,
                                tol=tol)
        theta = torch.tensor(np.random.default_rng(1).standard_normal(prob.m), device=dev,
                             requires_grad=True)
        x = AG.solve(torch.nn.functional.softplus(theta), torch.tensor(b0, device=dev), prob)
        ((x - torch.tensor(xs, device=dev)) ** 2).sum().backward()
        grads[prec] = theta.grad.cpu().numpy()
    g32, g64 = grads[rchol_amd.FP32], grads[rchol_amd.FP64]
    err = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"100^3 gradient: mixed vs FP64 adjoint {err:.1e}", flush=True)
    assert err <= 1e-6
