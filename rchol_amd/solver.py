"""Device-resident PCG solver over the C ABI of librchol_amd.so.

Python face of the hot path of the reference (c++/util/pcg.cpp and
python/rchol/util.py:32-61).  A ``Solver`` owns one engine context: A and the
factor G are uploaded once and stay resident in HBM across solves.
"""
import ctypes
import hashlib

import numpy as np
import scipy.sparse as sp

from . import _lib

FP64 = 0
FP32 = 1

OPT_GRAPH_ITERS = 1
OPT_RESID_REPLACE = 2
OPT_SPMV_LANES = 3
OPT_TRSV_LANES = 4
OPT_PROFILE = 5
OPT_SERIAL_WIDTH = 6
OPT_COHERENT_GATHER = 7
OPT_TASK_WAVES = 8
OPT_STRICT_SYNC = 9
OPT_CHUNKED = 10
OPT_WAIT_TICKS = 11


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_size_t),
        ("nnz_a", ctypes.c_size_t),
        ("nnz_g", ctypes.c_size_t),
        ("fwd_levels", ctypes.c_int),
        ("bwd_levels", ctypes.c_int),
        ("fwd_tasks", ctypes.c_int),
        ("bwd_tasks", ctypes.c_int),
        ("trsv_waves", ctypes.c_int),
        ("g_precision", ctypes.c_int),
        ("last_iters", ctypes.c_int),
        ("last_solve_ms", ctypes.c_double),
        ("last_relres", ctypes.c_double),
        ("kernel_ms", ctypes.c_double * 4),
        ("kernel_iters", ctypes.c_int),
        ("setup_ms", ctypes.c_double * 4),
        ("chunked", ctypes.c_int),
        ("chunks", ctypes.c_int * 2),
        ("chunk_rounds", ctypes.c_longlong * 2),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["kernel_ms"] = list(self.kernel_ms)
        d["setup_ms"] = list(self.setup_ms)
        d["chunks"] = list(self.chunks)
        d["chunk_rounds"] = list(self.chunk_rounds)
        return d


def _pattern_key(A):
    """Digest of a CSR matrix's shape, indptr and indices."""
    h = hashlib.blake2b(digest_size=16)
    h.update(np.asarray(A.shape, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(A.indptr, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(A.indices, dtype=np.int64).tobytes())
    return h.digest()


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(_lib.u64p)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_lib.f64p)


def device_count():
    return _lib.engine().rchol_amd_device_count()


def stored_from_lower(G):
    """Stored factor arrays (CSC of lower G == CSR of G^T) from a scipy G."""
    Gc = sp.csc_matrix(G)
    Gc.sort_indices()
    return Gc.indptr, Gc.indices, Gc.data


class Solver:
    """One engine context: A and G resident on one GPU.

    A: scipy sparse (n x n), used as CSR.
    G: either the stored triple (colPtr, rowIdx, val) exactly as the C++
       rchol() fills SparseCSR G, or a scipy lower-triangular matrix.
    precision: FP64 or FP32 for the factor and the two triangular sweeps.
    """

    def __init__(self, A=None, G=None, precision=FP64, device=0, graph_iters=None,
                 spmv_lanes=None, trsv_lanes=None, chunked=None):
        self._lib = _lib.engine()
        h = ctypes.c_void_p()
        _lib.check(self._lib.rchol_amd_create(int(device), ctypes.byref(h)))
        self._h = h
        self.n = None
        if graph_iters is not None:
            self.set_option(OPT_GRAPH_ITERS, graph_iters)
        if spmv_lanes is not None:
            self.set_option(OPT_SPMV_LANES, spmv_lanes)
        if trsv_lanes is not None:
            self.set_option(OPT_TRSV_LANES, trsv_lanes)
        if chunked is not None:
            self.set_option(OPT_CHUNKED, chunked)
        if A is not None:
            self.upload_A(A)
        if G is not None:
            self.upload_G(G, precision)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.rchol_amd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, key, value):
        _lib.check(self._lib.rchol_amd_set_option(self._h, int(key), float(value)))

    def upload_A(self, A):
        A = sp.csr_matrix(A)
        A.sort_indices()
        n = A.shape[0]
        rp, prp = _u64(A.indptr)
        ci, pci = _u64(A.indices)
        v, pv = _f64(A.data)
        _lib.check(self._lib.rchol_amd_upload_A(self._h, n, prp, pci, pv))
        self.n = n
        self._pattern = _pattern_key(A)

    def update_A_values(self, A, check_pattern=True):
        """New values for the resident matrix, whose sparsity pattern is
        unchanged (the ex_reuse_partition.py sequence): one copy of the
        values.  check_pattern compares a digest of indptr/indices with the
        uploaded matrix's and raises ValueError on a different pattern."""
        A = sp.csr_matrix(A)
        A.sort_indices()
        if check_pattern and _pattern_key(A) != getattr(self, "_pattern", None):
            raise ValueError("sparsity pattern differs from the resident matrix")
        v, pv = _f64(A.data)
        _lib.check(self._lib.rchol_amd_update_A_values(self._h, v.size, pv))

    def upload_G(self, G, precision=FP64):
        if isinstance(G, tuple):
            cp, ri, gv = G
        else:
            cp, ri, gv = stored_from_lower(G)
        cp, pcp = _u64(cp)
        ri, pri = _u64(ri)
        gv, pgv = _f64(gv)
        n = cp.size - 1
        _lib.check(self._lib.rchol_amd_upload_G(self._h, n, pcp, pri, pgv, int(precision)))
        if self.n is None:
            self.n = n

    def copy_probe(self, nbytes=1 << 30, reps=5):
        """Achievable HBM GB/s of a plain streaming copy (read + write)."""
        g, m = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._lib.rchol_amd_copy_probe(self._h, int(nbytes), int(reps), ctypes.byref(g),
                                                  ctypes.byref(m)))
        return {"gbs": g.value, "ms": m.value, "bytes": int(nbytes)}

    def autotune(self, reps=3):
        """Pick lanes-per-row of the sweeps and the SpMV by timing them."""
        tl, sl = ctypes.c_int(), ctypes.c_int()
        tm, sm = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._lib.rchol_amd_autotune(self._h, int(reps), ctypes.byref(tl),
                                                ctypes.byref(sl), ctypes.byref(tm),
                                                ctypes.byref(sm)))
        return {"trsv_lanes": tl.value, "spmv_lanes": sl.value, "precond_ms": tm.value,
                "spmv_ms": sm.value}

    def set_segments(self, off):
        """Uneven batch: A and G are block-diagonal stacks, system s owns rows
        off[s]..off[s+1]-1."""
        o, po = _u64(off)
        _lib.check(self._lib.rchol_amd_set_segments(self._h, o.size - 1, po))
        self.nseg = o.size - 1

    def pcg_segments(self, b, tol, maxit, x0=None):
        """One PCG per stacked system -> (x, relres[nseg], itr[nseg])."""
        b, pb = _f64(b)
        keep = 0
        x = np.zeros_like(b)
        if x0 is not None:
            x[:] = x0
            keep = 1
        rr = np.zeros(self.nseg)
        it = np.zeros(self.nseg, dtype=np.int32)
        _lib.check(self._lib.rchol_amd_pcg_segments(self._h, pb, float(tol), int(maxit),
                                                    x.ctypes.data_as(_lib.f64p), keep,
                                                    rr.ctypes.data_as(_lib.f64p),
                                                    it.ctypes.data_as(_lib.i32p)))
        return x, rr, it

    def set_perm(self, perm):
        """Vectors of later solves are in the original numbering; A and G
        factor A(perm, perm) (perm=None clears)."""
        if perm is None:
            _lib.check(self._lib.rchol_amd_set_perm(self._h, 0, None))
            return
        p, pp = _u64(perm)
        _lib.check(self._lib.rchol_amd_set_perm(self._h, p.size, pp))

    def edge_grad_device(self, m, u_ptr, v_ptr, lam_ptr, x_ptr, g_ptr):
        _lib.check(self._lib.rchol_amd_edge_grad_device(
            self._h, int(m), ctypes.c_void_p(u_ptr), ctypes.c_void_p(v_ptr),
            ctypes.c_void_p(lam_ptr), ctypes.c_void_p(x_ptr), ctypes.c_void_p(g_ptr)))

    def precond(self, r):
        """z = G^-T G^-1 r (pcg.cpp:141-159)."""
        r, pr = _f64(r)
        z = np.empty_like(r)
        _lib.check(self._lib.rchol_amd_precond(self._h, pr, z.ctypes.data_as(_lib.f64p)))
        return z

    def spmv(self, x):
        x, px = _f64(x)
        y = np.empty_like(x)
        _lib.check(self._lib.rchol_amd_spmv(self._h, px, y.ctypes.data_as(_lib.f64p)))
        return y

    def pcg(self, b, tol, maxit, x0=None):
        """PCG of pcg.cpp:57-127 on host arrays -> (x, relres, itr)."""
        b, pb = _f64(b)
        if x0 is None:
            x = np.zeros_like(b)
            keep = 0
        else:
            x = np.array(x0, dtype=np.float64, copy=True)
            keep = 1
        relres = ctypes.c_double()
        itr = ctypes.c_int()
        _lib.check(self._lib.rchol_amd_pcg(self._h, pb, float(tol), int(maxit),
                                           x.ctypes.data_as(_lib.f64p), keep,
                                           ctypes.byref(relres), ctypes.byref(itr)))
        return x, relres.value, itr.value

    def pcg_batch(self, B, tol, maxit, X0=None):
        """k independent right-hand sides (columns of B, n x k, k <= 32) solved
        together -> (X, relres[k], itr[k])."""
        B = np.ascontiguousarray(B, dtype=np.float64)
        n, k = B.shape
        keep = 0
        X = np.zeros_like(B)
        if X0 is not None:
            X[:] = X0
            keep = 1
        rr = np.zeros(k)
        it = np.zeros(k, dtype=np.int32)
        _lib.check(self._lib.rchol_amd_pcg_batch(self._h, int(k), B.ctypes.data_as(_lib.f64p),
                                                 float(tol), int(maxit),
                                                 X.ctypes.data_as(_lib.f64p), keep,
                                                 rr.ctypes.data_as(_lib.f64p),
                                                 it.ctypes.data_as(_lib.i32p)))
        return X, rr, it

    def precond_batch(self, R):
        R = np.ascontiguousarray(R, dtype=np.float64)
        n, k = R.shape
        Z = np.empty_like(R)
        _lib.check(self._lib.rchol_amd_precond_batch(self._h, int(k), R.ctypes.data_as(_lib.f64p),
                                                     Z.ctypes.data_as(_lib.f64p)))
        return Z

    def pcg_batch_device(self, k, b_ptr, x_ptr, tol, maxit, x_is_initial=False):
        rr = np.zeros(k)
        it = np.zeros(k, dtype=np.int32)
        _lib.check(self._lib.rchol_amd_pcg_batch_device(self._h, int(k), ctypes.c_void_p(b_ptr),
                                                        float(tol), int(maxit),
                                                        ctypes.c_void_p(x_ptr),
                                                        int(bool(x_is_initial)),
                                                        rr.ctypes.data_as(_lib.f64p),
                                                        it.ctypes.data_as(_lib.i32p)))
        return rr, it

    def pcg_device(self, b_ptr, x_ptr, tol, maxit, x_is_initial=False):
        """PCG on device pointers (e.g. torch tensors' data_ptr())."""
        relres = ctypes.c_double()
        itr = ctypes.c_int()
        _lib.check(self._lib.rchol_amd_pcg_device(self._h, ctypes.c_void_p(b_ptr), float(tol),
                                                  int(maxit), ctypes.c_void_p(x_ptr),
                                                  int(bool(x_is_initial)),
                                                  ctypes.byref(relres), ctypes.byref(itr)))
        return relres.value, itr.value

    def precond_device(self, r_ptr, z_ptr):
        _lib.check(self._lib.rchol_amd_precond_device(self._h, ctypes.c_void_p(r_ptr),
                                                      ctypes.c_void_p(z_ptr)))

    def precond_device_noalloc(self, r_ptr, z_ptr):
        """precond_device with no device allocation: the engine's static
        64 MiB scratch per GPU (calls on one GPU must not overlap)."""
        _lib.check(self._lib.rchol_amd_precond_device_noalloc(self._h, ctypes.c_void_p(r_ptr),
                                                              ctypes.c_void_p(z_ptr)))

    def stats(self):
        s = Stats()
        _lib.check(self._lib.rchol_amd_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()


def stack_systems(systems):
    """Block-diagonal stack of [(A_s, (colPtr, rowIdx, val)_s)] -> (A, G, off)."""
    As, cps, ris, vs, off = [], [], [0], [], [0]
    cps = [np.zeros(1, dtype=np.uint64)]
    ris, vs = [], []
    base_n, base_nnz = 0, 0
    for A, (cp, ri, v) in systems:
        As.append(sp.csr_matrix(A))
        cp = np.asarray(cp, dtype=np.uint64)
        cps.append(cp[1:] + np.uint64(base_nnz))
        ris.append(np.asarray(ri, dtype=np.uint64) + np.uint64(base_n))
        vs.append(np.asarray(v, dtype=np.float64))
        base_n += A.shape[0]
        base_nnz += int(cp[-1])
        off.append(base_n)
    A = sp.block_diag(As, format="csr")
    G = (np.concatenate(cps), np.concatenate(ris), np.concatenate(vs))
    return A, G, np.asarray(off, dtype=np.uint64)


def analyze(G, serial_width=0):
    """Host-only level schedule of the two sweeps for a factor (stored triple
    or scipy lower G): dict per sweep with levels, split, slot order and the
    per-slot proxies of the persistent sweep's waits."""
    if isinstance(G, tuple):
        cp, ri, gv = G
    else:
        cp, ri, gv = stored_from_lower(G)
    cp, pcp = _u64(cp)
    ri, pri = _u64(ri)
    gv, pgv = _f64(gv)
    n = cp.size - 1
    sc = _lib.Schedule()
    lib = _lib.engine()
    _lib.check(lib.rchol_amd_analyze_G(n, pcp, pri, pgv, int(serial_width), ctypes.byref(sc)))
    out = {}
    for k, name in enumerate(("fwd", "bwd")):
        nl = sc.nlevels[k]
        out[name] = {
            "nlevels": nl, "head_end": sc.head_end[k], "tail_begin": sc.tail_begin[k],
            "mid_tasks": sc.mid_tasks[k],
            "level_ptr": np.ctypeslib.as_array(sc.level_ptr[k], shape=(nl + 1,)).copy(),
            "slot_row": np.ctypeslib.as_array(sc.slot_row[k], shape=(n,)).copy(),
            "proxy": np.ctypeslib.as_array(sc.proxy[k], shape=(n,)).copy(),
        }
    lib.rchol_amd_schedule_free(ctypes.byref(sc))
    return out


def chunk_apply_host(G, r, max_rows=65535):
    """The chunked sweeps' plan for G (chunks of at most max_rows rows) run
    on the host in the device kernel's order of operations: z = G^-T G^-1 r.
    Raises on a plan that breaks its invariants.  Returns (z, [chunks,
    forward parts, backward parts], rounds of both sweeps, critical-path
    estimate of [forward, backward] in microseconds)."""
    if isinstance(G, tuple):
        cp, ri, gv = G
    else:
        cp, ri, gv = stored_from_lower(G)
    cp, pcp = _u64(cp)
    ri, pri = _u64(ri)
    gv, pgv = _f64(gv)
    r, pr = _f64(r)
    z = np.empty_like(r)
    cnt = np.zeros(3, dtype=np.int32)
    nrd = ctypes.c_longlong()
    est = np.zeros(2)
    lib = _lib.engine()
    _lib.check(lib.rchol_amd_chunk_apply_host(cp.size - 1, pcp, pri, pgv, int(max_rows), pr,
                                              z.ctypes.data_as(_lib.f64p),
                                              cnt.ctypes.data_as(_lib.i32p), ctypes.byref(nrd),
                                              est.ctypes.data_as(_lib.f64p)))
    return z, cnt.tolist(), nrd.value, est


def pcg(A, b, epsilon, maxitr, G, Gt=None):
    """Drop-in for python/rchol/util.py:32-61: pcg(A, b, epsilon, maxitr, G, Gt).

    G is the lower-triangular factor (scipy), Gt its transpose (accepted for
    signature compatibility; the engine derives both sweeps from G).
    Returns (x, relres, niters) with relres the true residual
    ||A x - b|| / ||b|| (util.py:56).  The preconditioner applied is
    G^-T G^-1 as in c++/util/pcg.cpp:141-159.
    """
    with Solver(A, G) as s:
        return s.pcg(b, epsilon, maxitr)
