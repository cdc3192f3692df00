"""Host factorization, ordering and problem generators (librchol_factor.so).

Python face of the producer side of the solve path:
  rchol(A, nt=1, perm, part) -> (G, perm, part)     python/rchol/rchol.py:8-45
  laplace_3d(n)                                     c++/util/laplace_3d.hpp:8-65
  nd_graph / nd_grid                                find_separator.cpp:109-163 contract
G is returned as the scipy CSR of the lower-triangular factor (as rchol.py
does) and, through factor(), as the stored triple of the C++ SparseCSR G.
"""
import ctypes
import math
import weakref

import numpy as np
import scipy.sparse as sp

from . import _lib

ORDER_STATIC = 0
ORDER_MIS = 1
_ORDERS = {"static": ORDER_STATIC, "mis": ORDER_MIS}


def _chk(rc):
    if rc != 0:
        msg = _lib.factor().rchol_factor_last_error()
        raise RuntimeError(f"rchol_factor error {rc}: {msg.decode() if msg else ''}")


def _take_csr(n, rp, ci, v):
    N = n.value
    rpa = _owned(rp, N + 1)
    nnz = int(rpa[-1])
    # indices below 2^63: the int64 views share the buffers (scipy keeps or
    # narrows them; the values are not copied)
    return sp.csr_matrix((_owned(v, nnz), _owned(ci, nnz).view(np.int64), rpa.view(np.int64)),
                         shape=(N, N))


def laplace_3d(side):
    """7-point Dirichlet Laplacian on side^3 (laplace_3d.hpp:8-65)."""
    n, rp, ci, v = ctypes.c_size_t(), _lib.u64p(), _lib.u64p(), _lib.f64p()
    _chk(_lib.factor().rchol_laplace_3d(int(side), ctypes.byref(n), ctypes.byref(rp),
                                        ctypes.byref(ci), ctypes.byref(v)))
    return _take_csr(n, rp, ci, v)


def varcoef_3d(nx, ny, nz, sigma=1.5, shift=1e-3, seed=0):
    """7-point SDDM with face conductivities exp(N(0, sigma^2)) + shift*I."""
    n, rp, ci, v = ctypes.c_size_t(), _lib.u64p(), _lib.u64p(), _lib.f64p()
    _chk(_lib.factor().rchol_varcoef_3d(int(nx), int(ny), int(nz), float(sigma), float(shift),
                                        int(seed), ctypes.byref(n), ctypes.byref(rp),
                                        ctypes.byref(ci), ctypes.byref(v)))
    return _take_csr(n, rp, ci, v)


def grid_edges(nx, ny, nz):
    m, u, v = ctypes.c_size_t(), _lib.u64p(), _lib.u64p()
    _chk(_lib.factor().rchol_grid_edges(int(nx), int(ny), int(nz), ctypes.byref(m),
                                        ctypes.byref(u), ctypes.byref(v)))
    M = m.value
    ua = np.ctypeslib.as_array(u, shape=(max(M, 1),))[:M].astype(np.int64)
    va = np.ctypeslib.as_array(v, shape=(max(M, 1),))[:M].astype(np.int64)
    L = _lib.factor()
    L.rchol_factor_free(ctypes.cast(u, ctypes.c_void_p))
    L.rchol_factor_free(ctypes.cast(v, ctypes.c_void_p))
    return ua, va


def grid_laplacian(nx, ny, nz, w, shift):
    """Weighted 6-neighbour grid Laplacian (edges in grid_edges order) + shift*I."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    n, rp, ci, v = ctypes.c_size_t(), _lib.u64p(), _lib.u64p(), _lib.f64p()
    _chk(_lib.factor().rchol_grid_laplacian(int(nx), int(ny), int(nz),
                                            w.ctypes.data_as(_lib.f64p), float(shift),
                                            ctypes.byref(n), ctypes.byref(rp), ctypes.byref(ci),
                                            ctypes.byref(v)))
    return _take_csr(n, rp, ci, v)


def permute(A, perm):
    """A(perm, perm) with sorted rows (util.cpp:16-56 reorder), threaded."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, dtype=np.uint64)
    ci = np.ascontiguousarray(A.indices, dtype=np.uint64)
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    p = np.ascontiguousarray(perm, dtype=np.uint64)
    orp, oci, ov = _lib.u64p(), _lib.u64p(), _lib.f64p()
    _chk(_lib.factor().rchol_permute_csr(n, rp.ctypes.data_as(_lib.u64p),
                                         ci.ctypes.data_as(_lib.u64p), v.ctypes.data_as(_lib.f64p),
                                         p.ctypes.data_as(_lib.u64p), ctypes.byref(orp),
                                         ctypes.byref(oci), ctypes.byref(ov)))
    return _take_csr(ctypes.c_size_t(n), orp, oci, ov)


def nd_grid(nx, ny, nz, depth):
    """Geometric nested dissection -> (perm, block sizes in post-order)."""
    n = int(nx) * int(ny) * int(nz)
    perm = np.empty(n, dtype=np.uint64)
    blocks = np.empty((1 << depth) - 1, dtype=np.uint64)
    _chk(_lib.factor().rchol_nd_grid(int(nx), int(ny), int(nz), int(depth),
                                     perm.ctypes.data_as(_lib.u64p),
                                     blocks.ctypes.data_as(_lib.u64p)))
    return perm, blocks


def nd_graph(A, depth):
    """Nested dissection of the graph of A (find_separator contract)."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, dtype=np.uint64)
    ci = np.ascontiguousarray(A.indices, dtype=np.uint64)
    perm = np.empty(n, dtype=np.uint64)
    blocks = np.empty((1 << depth) - 1, dtype=np.uint64)
    rc = _lib.factor().rchol_nd_graph(n, rp.ctypes.data_as(_lib.u64p),
                                      ci.ctypes.data_as(_lib.u64p), int(depth),
                                      perm.ctypes.data_as(_lib.u64p),
                                      blocks.ctypes.data_as(_lib.u64p))
    if rc == -3:
        raise Exception("too many threads requested")  # find_separator.cpp:125
    _chk(rc)
    return perm, blocks


def factor(A, perm=None, blocks=None, order="static", seed=0):
    """Randomized Cholesky of SDDM A.  Returns ((colPtr, rowIdx, val), perm, seconds)
    with the stored arrays of the C++ SparseCSR G for A(perm, perm)."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, dtype=np.uint64)
    ci = np.ascontiguousarray(A.indices, dtype=np.uint64)
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    if blocks is None:
        blocks = np.array([n], dtype=np.uint64)
    blocks = np.ascontiguousarray(blocks, dtype=np.uint64)
    pp = None
    if perm is not None:
        perm = np.ascontiguousarray(perm, dtype=np.uint64)
        pp = perm.ctypes.data_as(_lib.u64p)
    out = _lib.FactorOut()
    _chk(_lib.factor().rchol_factor(n, rp.ctypes.data_as(_lib.u64p), ci.ctypes.data_as(_lib.u64p),
                                    v.ctypes.data_as(_lib.f64p), pp,
                                    blocks.ctypes.data_as(_lib.u64p), blocks.size,
                                    _ORDERS[order], int(seed) & 0xFFFFFFFF, ctypes.byref(out)))
    cp = _owned(out.colPtr, n + 1)
    nnz = int(cp[-1])
    ri = _owned(out.rowIdx, nnz)
    gv = _owned(out.val, nnz)
    po = _owned(out.perm, n)
    return (cp, ri, gv), po, out.seconds


def _owned(ptr, count):
    """numpy view of count elements of a buffer the factorization malloc'ed,
    released by rchol_factor_free once the last view of it is gone (no copy:
    the factor's arrays run to gigabytes)."""
    base = np.ctypeslib.as_array(ptr, shape=(max(count, 1),))
    weakref.finalize(base, _lib.factor().rchol_factor_free, ctypes.cast(ptr, ctypes.c_void_p))
    return base[:count]


def lower(Gs):
    """scipy CSR of lower G from the stored triple."""
    cp, ri, v = Gs
    n = cp.size - 1
    return sp.csc_matrix((v, ri.astype(np.int64), cp.astype(np.int64)), shape=(n, n)).tocsr()


def rchol(A, nt=1, perm=np.array([]), part=np.array([]), seed=0, order="static", grid=None):
    """Drop-in for python/rchol/rchol.py:8-45: returns (G, perm, part), G lower CSR
    with G G^T ~ A[perm][:, perm].

    nt: power-of-two thread count (the ND tree has nt leaves); perm/part may
    be given together to reuse an ordering (ex_reuse_partition.py).  grid =
    (nx, ny, nz) selects the geometric ND for grid-numbered matrices; other
    matrices use the breadth-first-level ND.  order="mis" reorders inside the
    ND blocks by independent-set layers (shallow triangular sweeps)."""
    if nt < 1:
        raise Exception("number of threads must be at least 1")
    perm = np.asarray(perm)
    part = np.asarray(part)
    if (perm.size > 0) ^ (part.size > 0):
        raise Exception("permutation and partition must be given together")
    if not math.log2(nt).is_integer():
        nt = 2 ** math.floor(math.log2(nt))
        print('number of threads rounded down to {}'.format(nt))
    n = A.shape[0]
    if perm.size == 0:
        if nt == 1:
            perm = np.arange(n, dtype=np.uint64)
            part = np.array([0, n + 1], dtype=np.uint64)
        else:
            depth = int(math.log2(nt)) + 1
            if grid is not None:
                perm, val = nd_grid(grid[0], grid[1], grid[2], depth)
            else:
                perm, val = nd_graph(A, depth)
            part = np.append(0, np.cumsum(val)).astype(np.uint64)
            part[-1] = part[-1] + 1
    blocks = np.diff(np.asarray(part, dtype=np.int64))
    blocks[-1] -= 1
    seq = blocks.size == 1 and perm.size == n and np.array_equal(perm, np.arange(n))
    Gs, perm_out, _ = factor(A, None if seq else perm, blocks, order, seed)
    return lower(Gs), perm_out, np.asarray(part, dtype=np.uint64)
