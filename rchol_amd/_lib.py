"""ctypes loader for the in-tree shared libraries.

librchol_amd.so    -- the HIP solve engine (include/rchol_amd.h)
librchol_factor.so -- host factorization and ordering (include/rchol_factor.h)

Both are built in-tree by ``make -C rchol_amd/csrc`` (``__graft_entry__.build``).
A missing library raises immediately: there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_PATH = os.path.join(_HERE, "librchol_amd.so")
FACTOR_PATH = os.path.join(_HERE, "librchol_factor.so")

u64p = ctypes.POINTER(ctypes.c_uint64)
f64p = ctypes.POINTER(ctypes.c_double)
i32p = ctypes.POINTER(ctypes.c_int)

class Schedule(ctypes.Structure):
    _fields_ = [("nlevels", ctypes.c_int * 2), ("head_end", ctypes.c_int * 2),
                ("tail_begin", ctypes.c_int * 2), ("mid_tasks", ctypes.c_int * 2),
                ("level_ptr", i32p * 2), ("slot_row", i32p * 2),
                ("proxy", ctypes.POINTER(ctypes.c_uint16) * 2)]


# Every C-ABI symbol of include/rchol_amd.h with its signature.
ENGINE_SYMBOLS = {
    "rchol_amd_analyze_G": (ctypes.c_int, [ctypes.c_size_t, u64p, u64p, f64p, ctypes.c_int,
                                           ctypes.POINTER(Schedule)]),
    "rchol_amd_schedule_free": (None, [ctypes.POINTER(Schedule)]),
    "rchol_amd_last_error": (ctypes.c_char_p, []),
    "rchol_amd_device_count": (ctypes.c_int, []),
    "rchol_amd_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "rchol_amd_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "rchol_amd_upload_A": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, u64p, u64p, f64p]),
    "rchol_amd_update_A_values": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, f64p]),
    "rchol_amd_upload_G": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, u64p, u64p, f64p,
                                          ctypes.c_int]),
    "rchol_amd_precond": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p]),
    "rchol_amd_precond_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "rchol_amd_chunk_apply_host": (ctypes.c_int, [ctypes.c_size_t, u64p, u64p, f64p, ctypes.c_int,
                                                  f64p, f64p, i32p,
                                                  ctypes.POINTER(ctypes.c_longlong), f64p]),
    "rchol_amd_precond_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "rchol_amd_spmv": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p]),
    "rchol_amd_pcg": (ctypes.c_int, [ctypes.c_void_p, f64p, ctypes.c_double, ctypes.c_int, f64p,
                                     ctypes.c_int, f64p, i32p]),
    "rchol_amd_pcg_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_void_p, ctypes.c_int, f64p,
                                            i32p]),
    "rchol_amd_pcg_csr": (ctypes.c_int, [ctypes.c_size_t, u64p, u64p, f64p, f64p,
                                         ctypes.c_double, ctypes.c_int, u64p, u64p, f64p, f64p,
                                         ctypes.c_int, f64p, i32p]),
    "rchol_amd_pcg_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, f64p, ctypes.c_double,
                                           ctypes.c_int, f64p, ctypes.c_int, f64p, i32p]),
    "rchol_amd_pcg_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                  ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                                  ctypes.c_int, f64p, i32p]),
    "rchol_amd_precond_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, f64p, f64p]),
    "rchol_amd_set_perm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, u64p]),
    "rchol_amd_edge_grad_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t,
                                                  ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    "rchol_amd_set_segments": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, u64p]),
    "rchol_amd_pcg_segments": (ctypes.c_int, [ctypes.c_void_p, f64p, ctypes.c_double, ctypes.c_int,
                                              f64p, ctypes.c_int, f64p, i32p]),
    "rchol_amd_copy_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, f64p,
                                            f64p]),
    "rchol_amd_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, i32p, i32p, f64p, f64p]),
    "rchol_amd_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    "rchol_amd_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
}

size_tp = ctypes.POINTER(ctypes.c_size_t)
pu64p = ctypes.POINTER(u64p)
pf64p = ctypes.POINTER(f64p)


class FactorOut(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("colPtr", u64p), ("rowIdx", u64p), ("val", f64p),
                ("perm", u64p), ("seconds", ctypes.c_double)]


# Every C-ABI symbol of include/rchol_factor.h with its signature.
FACTOR_SYMBOLS = {
    "rchol_factor_last_error": (ctypes.c_char_p, []),
    "rchol_factor_free": (None, [ctypes.c_void_p]),
    "rchol_factor": (ctypes.c_int, [ctypes.c_size_t, u64p, u64p, f64p, u64p, u64p,
                                    ctypes.c_size_t, ctypes.c_int, ctypes.c_uint,
                                    ctypes.POINTER(FactorOut)]),
    "rchol_nd_grid": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     u64p, u64p]),
    "rchol_nd_graph": (ctypes.c_int, [ctypes.c_size_t, u64p, u64p, ctypes.c_int, u64p, u64p]),
    "rchol_laplace_3d": (ctypes.c_int, [ctypes.c_int, size_tp, pu64p, pu64p, pf64p]),
    "rchol_varcoef_3d": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong,
                                        size_tp, pu64p, pu64p, pf64p]),
    "rchol_grid_laplacian": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p,
                                            ctypes.c_double, size_tp, pu64p, pu64p, pf64p]),
    "rchol_grid_edges": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, size_tp,
                                        pu64p, pu64p]),
    "rchol_permute_csr": (ctypes.c_int, [ctypes.c_size_t, u64p, u64p, f64p, u64p, pu64p, pu64p,
                                         pf64p]),
}

_engine = None
_factor = None


def _bind(lib, table):
    for name, (res, args) in table.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def _share_torch_runtime():
    """Import PyTorch, when it is installed, before the engine library.

    PyTorch-ROCm ships its own HIP/HSA runtime, which it requests under the
    unversioned name libamdhip64.so.  The engine links libamdhip64.so.7: loaded
    after torch it binds to torch's copy (same SONAME) and one runtime serves
    both; loaded before torch, torch would start a second runtime in the
    process, which finds no GPU.  The engine itself does not use torch."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def engine():
    """The HIP engine library (loaded once)."""
    global _engine
    if _engine is None:
        _share_torch_runtime()
        if not os.path.exists(ENGINE_PATH):
            raise RuntimeError(
                f"{ENGINE_PATH} is not built; run `make -C rchol_amd/csrc` "
                "(or __graft_entry__.build()) -- there is no CPU fallback")
        lib = ctypes.CDLL(ENGINE_PATH)
        _bind(lib, ENGINE_SYMBOLS)
        _engine = lib
    return _engine


def factor():
    """The host factorization library (loaded once)."""
    global _factor
    if _factor is None:
        if not os.path.exists(FACTOR_PATH):
            raise RuntimeError(f"{FACTOR_PATH} is not built; run `make -C rchol_amd/csrc`")
        lib = ctypes.CDLL(FACTOR_PATH)
        _bind(lib, FACTOR_SYMBOLS)
        _factor = lib
    return _factor


def check(rc, lib=None):
    if rc != 0:
        lib = lib or engine()
        msg = lib.rchol_amd_last_error()
        raise RuntimeError(f"rchol_amd error {rc}: {msg.decode() if msg else ''}")
