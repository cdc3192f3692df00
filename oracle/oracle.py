"""ORACLE -- test infrastructure only.

Only tests/, the smoke check in __graft_entry__.py and the CPU-baseline leg of
bench.py may import this module, and only as the checker: nothing in the
product (rchol_amd/) imports it.

Python access to
  * liboracle.so           -- the C restatement of c++/util/pcg.cpp
                              (pcg_oracle.c; pcg.cpp:57-159);
  * _ref/librchol_ref.so   -- the reference's own host factorization
                              (rchol.cpp, rchol_lap.cpp, spcol.cpp, util.cpp,
                              sparse.cpp), built from the reference sources by
                              `make -C oracle ref` with a seed hook, present
                              only where it was built.
Matrices are scipy CSR with the reference's 64-bit index arrays.  A factor is
passed as the "stored" triple (colPtr, rowIdx, val) exactly as rchol() fills
SparseCSR G: CSC of lower G == CSR of upper G^T (rchol_lap.cpp:358-387).
"""
import ctypes
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_lib = None
_ref = None


def _p64(a):
    return a.ctypes.data_as(_u64p)


def _pf(a):
    return a.ctypes.data_as(_f64p)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        for name in ("oracle_spmv", "oracle_fwd", "oracle_bwd", "oracle_precond"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_size_t, _u64p, _u64p, _f64p, _f64p, _f64p]
        L.oracle_pcg.restype = ctypes.c_int
        L.oracle_pcg.argtypes = [ctypes.c_size_t, _u64p, _u64p, _f64p, _f64p,
                                 ctypes.c_double, ctypes.c_int, _u64p, _u64p, _f64p,
                                 _f64p, _f64p, ctypes.POINTER(ctypes.c_int), _f64p]
        L.oracle_set_threads.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "librchol_ref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "librchol_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/librchol_ref.so is not built")
        L = ctypes.CDLL(path)
        pp64 = ctypes.POINTER(_u64p)
        ppf = ctypes.POINTER(_f64p)
        L.ref_rchol_seq.restype = ctypes.c_int
        L.ref_rchol_seq.argtypes = [ctypes.c_size_t, _u64p, _u64p, _f64p, ctypes.c_uint,
                                    pp64, pp64, ppf, ctypes.POINTER(ctypes.c_size_t)]
        L.ref_rchol_nd.restype = ctypes.c_int
        L.ref_rchol_nd.argtypes = [ctypes.c_size_t, _u64p, _u64p, _f64p, _u64p, _u64p,
                                   ctypes.c_size_t, ctypes.c_uint, pp64, pp64, ppf,
                                   ctypes.POINTER(ctypes.c_size_t)]
        L.ref_laplace_3d.restype = ctypes.c_int
        L.ref_laplace_3d.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t), pp64,
                                     pp64, ppf]
        L.ref_reorder.restype = ctypes.c_int
        L.ref_reorder.argtypes = [ctypes.c_size_t, _u64p, _u64p, _f64p, _u64p, pp64, pp64,
                                  ppf]
        L.ref_free.restype = None
        L.ref_free.argtypes = [ctypes.c_void_p]
        _ref = L
    return _ref


def set_threads(t):
    return lib().oracle_set_threads(int(t))


def _csr_arrays(A):
    A = sp.csr_matrix(A)
    rp = np.ascontiguousarray(A.indptr, dtype=np.uint64)
    ci = np.ascontiguousarray(A.indices, dtype=np.uint64)
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    return A.shape[0], rp, ci, v


def _stored(Gs):
    cp, ri, v = Gs
    return (np.ascontiguousarray(cp, dtype=np.uint64), np.ascontiguousarray(ri, dtype=np.uint64),
            np.ascontiguousarray(v, dtype=np.float64))


def spmv(A, x):
    n, rp, ci, v = _csr_arrays(A)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(n)
    lib().oracle_spmv(n, _p64(rp), _p64(ci), _pf(v), _pf(x), _pf(y))
    return y


def precond(Gs, r):
    """pcg.cpp:141-159: z = G^-T G^-1 r on the stored factor arrays."""
    cp, ri, v = _stored(Gs)
    n = cp.size - 1
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.empty(n)
    rc = lib().oracle_precond(n, _p64(cp), _p64(ri), _pf(v), _pf(r), _pf(z))
    if rc:
        raise ZeroDivisionError("zero pivot in G")
    return z


def fwd(Gs, r):
    cp, ri, v = _stored(Gs)
    n = cp.size - 1
    r = np.ascontiguousarray(r, dtype=np.float64)
    y = np.empty(n)
    if lib().oracle_fwd(n, _p64(cp), _p64(ri), _pf(v), _pf(r), _pf(y)):
        raise ZeroDivisionError("zero pivot in G")
    return y


def bwd(Gs, y):
    cp, ri, v = _stored(Gs)
    n = cp.size - 1
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.empty(n)
    if lib().oracle_bwd(n, _p64(cp), _p64(ri), _pf(v), _pf(y), _pf(z)):
        raise ZeroDivisionError("zero pivot in G")
    return z


def pcg(A, b, tol, maxit, Gs, x0=None, history=False):
    """pcg.cpp:57-127.  Returns (x, relres, itr[, ||r_k|| history])."""
    n, rp, ci, v = _csr_arrays(A)
    cp, ri, gv = _stored(Gs)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    relres = ctypes.c_double()
    itr = ctypes.c_int()
    hist = np.full(maxit + 1, np.nan)
    rc = lib().oracle_pcg(n, _p64(rp), _p64(ci), _pf(v), _pf(b), float(tol), int(maxit),
                          _p64(cp), _p64(ri), _pf(gv), _pf(x), ctypes.byref(relres),
                          ctypes.byref(itr), _pf(hist))
    if rc:
        raise ZeroDivisionError("zero pivot in G")
    if history:
        return x, relres.value, itr.value, hist[: itr.value + 1]
    return x, relres.value, itr.value


def _take_factor(L, cp, ri, gv, nnz, n):
    cpa = np.ctypeslib.as_array(cp, shape=(n + 1,)).copy()
    ria = np.ctypeslib.as_array(ri, shape=(max(nnz, 1),))[:nnz].copy()
    gva = np.ctypeslib.as_array(gv, shape=(max(nnz, 1),))[:nnz].copy()
    for p in (cp, ri, gv):
        L.ref_free(ctypes.cast(p, ctypes.c_void_p))
    return cpa, ria, gva


def ref_rchol_seq(A, seed):
    """Reference rchol(A, G) (rchol.cpp:7-36), seeded.  Returns stored arrays."""
    L = ref()
    n, rp, ci, v = _csr_arrays(A)
    cp, ri, gv = _u64p(), _u64p(), _f64p()
    nnz = ctypes.c_size_t()
    L.ref_rchol_seq(n, _p64(rp), _p64(ci), _pf(v), int(seed) & 0xFFFFFFFF, ctypes.byref(cp),
                    ctypes.byref(ri), ctypes.byref(gv), ctypes.byref(nnz))
    return _take_factor(L, cp, ri, gv, nnz.value, n)


def ref_rchol_nd(A, perm, blocks, seed):
    """Reference rchol_parallel.cpp:108-137 with a supplied ND (perm, block sizes)."""
    L = ref()
    n, rp, ci, v = _csr_arrays(A)
    perm = np.ascontiguousarray(perm, dtype=np.uint64)
    blocks = np.ascontiguousarray(blocks, dtype=np.uint64)
    cp, ri, gv = _u64p(), _u64p(), _f64p()
    nnz = ctypes.c_size_t()
    L.ref_rchol_nd(n, _p64(rp), _p64(ci), _pf(v), _p64(perm), _p64(blocks), blocks.size,
                   int(seed) & 0xFFFFFFFF, ctypes.byref(cp), ctypes.byref(ri), ctypes.byref(gv),
                   ctypes.byref(nnz))
    return _take_factor(L, cp, ri, gv, nnz.value, n)


def ref_laplace_3d(side):
    """Reference laplace_3d (laplace_3d.hpp:8-65) as scipy CSR."""
    L = ref()
    n = ctypes.c_size_t()
    rp, ci, v = _u64p(), _u64p(), _f64p()
    L.ref_laplace_3d(int(side), ctypes.byref(n), ctypes.byref(rp), ctypes.byref(ci),
                     ctypes.byref(v))
    N = n.value
    rpa = np.ctypeslib.as_array(rp, shape=(N + 1,)).copy()
    nnz = int(rpa[-1])
    cia = np.ctypeslib.as_array(ci, shape=(nnz,)).copy()
    va = np.ctypeslib.as_array(v, shape=(nnz,)).copy()
    for p in (rp, ci, v):
        L.ref_free(ctypes.cast(p, ctypes.c_void_p))
    return sp.csr_matrix((va, cia, rpa), shape=(N, N))


def ref_reorder(A, perm):
    """Reference reorder(A, P, B) (util.cpp:16-56)."""
    L = ref()
    n, rp, ci, v = _csr_arrays(A)
    perm = np.ascontiguousarray(perm, dtype=np.uint64)
    orp, oci, ov = _u64p(), _u64p(), _f64p()
    L.ref_reorder(n, _p64(rp), _p64(ci), _pf(v), _p64(perm), ctypes.byref(orp),
                  ctypes.byref(oci), ctypes.byref(ov))
    rpa = np.ctypeslib.as_array(orp, shape=(n + 1,)).copy()
    nnz = int(rpa[-1])
    cia = np.ctypeslib.as_array(oci, shape=(max(nnz, 1),))[:nnz].copy()
    va = np.ctypeslib.as_array(ov, shape=(max(nnz, 1),))[:nnz].copy()
    for p in (orp, oci, ov):
        L.ref_free(ctypes.cast(p, ctypes.c_void_p))
    return sp.csr_matrix((va, cia, rpa), shape=(n, n))


def stored_to_lower(Gs):
    """The stored arrays as the scipy CSR of lower-triangular G."""
    cp, ri, v = _stored(Gs)
    n = cp.size - 1
    return sp.csc_matrix((v, ri, cp), shape=(n, n)).tocsr()


class Prepared:
    """A and the stored factor converted once to the reference's 64-bit
    arrays, so that repeated oracle calls (the bench's CPU baseline) time the
    solve and not the conversion."""

    def __init__(self, A, Gs):
        self.n, self.rp, self.ci, self.v = _csr_arrays(A)
        self.cp, self.ri, self.gv = _stored(Gs)

    def pcg(self, b, tol, maxit):
        """pcg.cpp:57-127 from x0 = 0 -> (x, relres, itr)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        relres = ctypes.c_double()
        itr = ctypes.c_int()
        hist = np.empty(maxit + 1)
        rc = lib().oracle_pcg(self.n, _p64(self.rp), _p64(self.ci), _pf(self.v), _pf(b),
                              float(tol), int(maxit), _p64(self.cp), _p64(self.ri),
                              _pf(self.gv), _pf(x), ctypes.byref(relres), ctypes.byref(itr),
                              _pf(hist))
        if rc:
            raise ZeroDivisionError("zero pivot in G")
        return x, relres.value, itr.value

    def precond(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.n)
        if lib().oracle_precond(self.n, _p64(self.cp), _p64(self.ri), _pf(self.gv), _pf(r),
                                _pf(z)):
            raise ZeroDivisionError("zero pivot in G")
        return z
