"""Differentiable solve x = A(w)^-1 b for weighted grid Laplacians.

A(w) = sum_e w_e (e_u - e_v)(e_u - e_v)^T + shift * I over the edges of the
6-neighbour grid (rchol_amd.factor.grid_edges).  The forward pass factors
A(w) with rchol on the host (nested dissection, independent-set order) and
runs the device PCG with the FP32 factor sweeps; the backward pass runs the
adjoint PCG with the same factor (A is symmetric): lam = A^-1 dL/dx gives
dL/db = lam, and the edge-gradient kernel gives
dL/dw_e = -(lam_u - lam_v)(x_u - x_v).

PyTorch appears only here, for device memory and autograd plumbing; the
solve and the edge gradient are the engine's HIP kernels (C ABI).
"""
import hashlib

import numpy as np
import torch

from . import factor as F
from .solver import FP32, FP64, Solver


class GridLaplacian:
    """Problem description and factor cache for A(w) on an nx*ny*nz grid."""

    def __init__(self, nx, ny, nz, shift=1e-2, leaves=8, order="mis", precision=FP32, seed=1,
                 tol=1e-10, maxit=5000, device=0):
        self.dims = (int(nx), int(ny), int(nz))
        self.n = self.dims[0] * self.dims[1] * self.dims[2]
        self.shift = float(shift)
        self.depth = int(np.log2(leaves)) + 1
        self.order = order
        self.precision = precision
        self.seed = seed
        self.tol = tol
        self.maxit = maxit
        self.device = device
        u, v = F.grid_edges(*self.dims)
        self.m = u.size
        dev = torch.device("cuda", device)
        self.u = torch.as_tensor(u, dtype=torch.int64, device=dev)
        self.v = torch.as_tensor(v, dtype=torch.int64, device=dev)
        # the nested dissection depends on the grid only: computed once
        self._perm, self._blocks = F.nd_grid(*self.dims, self.depth)
        self._key = None
        self._solver = None
        self._w = None  # device copy of the weights the cached factor belongs to
        self.last = {}

    def matrix(self, w):
        return F.grid_laplacian(*self.dims, np.asarray(w, dtype=np.float64), self.shift)

    def solver(self, w):
        """Solver for A(w) (re-factored when w changes).

        The cache holds only the latest solver.  A replaced solver is not
        closed here: a forward pass whose backward is still pending keeps
        its own solver alive through its autograd context, and the context
        is freed (closing the solver) once that backward has run."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        key = hashlib.blake2b(w.tobytes(), digest_size=32).digest()
        if key != self._key:
            A = self.matrix(w)
            Gs, po, _ = F.factor(A, self._perm, self._blocks, self.order, self.seed)
            Ap = F.permute(A, po)
            s = Solver(Ap, Gs, self.precision, device=self.device)
            s.set_perm(po)
            self._solver = s
            self._key = key
        return self._solver

    def solver_for(self, w):
        """Solver for the weights of a device tensor: the cached factor is
        reused when w equals (on the device) the weights it was built for."""
        if self._solver is not None and self._w is not None and self._w.shape == w.shape \
                and torch.equal(self._w, w):
            return self._solver
        s = self.solver(w.detach().cpu().numpy())
        self._w = w.detach().clone()
        return s


class RcholSolve(torch.autograd.Function):
    """x = A(w)^-1 b with hand-written adjoint (see module docstring)."""

    @staticmethod
    def forward(ctx, w, b, problem):
        if w.dtype != torch.float64 or b.dtype != torch.float64:
            raise TypeError("w and b must be float64")
        if not (w.is_cuda and b.is_cuda):
            raise ValueError("w and b must live on the GPU")
        s = problem.solver_for(w.detach())
        b = b.contiguous()
        x = torch.empty_like(b)
        torch.cuda.current_stream(b.device).synchronize()
        relres, itr = s.pcg_device(b.data_ptr(), x.data_ptr(), problem.tol, problem.maxit)
        problem.last["forward"] = (relres, itr)
        ctx.problem = problem
        # the adjoint must use the factor of A(w) for this w, not whatever
        # the problem's cache holds when backward runs
        ctx.solver = s
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        problem = ctx.problem
        s = ctx.solver
        gx = gx.contiguous()
        lam = torch.empty_like(gx)
        torch.cuda.current_stream(gx.device).synchronize()
        relres, itr = s.pcg_device(gx.data_ptr(), lam.data_ptr(), problem.tol, problem.maxit)
        problem.last["adjoint"] = (relres, itr)
        gw = torch.empty(problem.m, dtype=torch.float64, device=gx.device)
        s.edge_grad_device(problem.m, problem.u.data_ptr(), problem.v.data_ptr(), lam.data_ptr(),
                           x.data_ptr(), gw.data_ptr())
        return gw, lam, None


def solve(w, b, problem):
    """Differentiable x = A(w)^-1 b."""
    return RcholSolve.apply(w, b, problem)


__all__ = ["GridLaplacian", "RcholSolve", "solve", "FP32", "FP64"]
