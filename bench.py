#!/usr/bin/env python3
"""PCG time-to-1e-10 on the 256^3 variable-coefficient Poisson problem,
mixed precision (FP32 factor sweeps) vs FP64, on one MI355X.

A step is one complete PCG solve (pcg.cpp:57-127 semantics) from x0 = 0 to
a relative residual of 1e-10 of the headline (mixed) variant, with A, G and
b already resident in HBM.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--side 256] [--leaves 128]

The CPU baseline is the oracle (C restatement of c++/util/pcg.cpp): one full
FP64 solve of the same system on 16 host threads (and on one core), which
also gives the iteration and preconditioner parity at full size.  The same
line reports a 32-right-hand-side batch on the headline factor.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(n, nnz_a, nnz_g, fp32):
    """Algorithmic (compulsory) HBM bytes per launch of the four loop kernels:
    each array is touched once; gathers of x are cache hits."""
    sv = 4 if fp32 else 8   # factor value bytes
    sx = 4 if fp32 else 8   # sweep vector bytes (y, z)
    fwd = nnz_g * (sv + 4) + 4 * (n + 1) + 8 * n + sx * n
    bwd = nnz_g * (sv + 4) + 4 * (n + 1) + sx * n + sx * n + 8 * n
    # p update (z, p read, p' written) + SpMV (A, p' and r read, q written)
    spmv = 12 * nnz_a + 4 * (n + 1) + sx * n + 8 * n + 8 * n + 16 * n + 8 * n
    upd = 32 * n + 16 * n + sx * n
    return [fwd, bwd, spmv, upd]


KERNEL_NAMES = ["sptrsv_fwd", "sptrsv_bwd", "spmv_p", "update"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--leaves", type=int, default=128)
    ap.add_argument("--order", default="mis", choices=["mis", "static"])
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--maxit", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle (CPU baseline, parity)")
    ap.add_argument("--no-cpu-one-core", action="store_true")
    ap.add_argument("--batch-cols", type=int, default=32,
                    help="right-hand sides of the batch run on the headline factor (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--no-static", action="store_true",
                    help="skip the FP64 solve on the reference's (static) order")
    ap.add_argument("--static-steps", type=int, default=2)
    ap.add_argument("--no-ref", action="store_true", help="skip the 1e-14 FP64 reference solve")
    ap.add_argument("--graph-iters", type=int, default=16)
    ap.add_argument("--trsv-lanes", type=int, default=1,
                    help="minimum lanes per sweep row (tasks widen to their longest row)")
    ap.add_argument("--spmv-lanes", type=int, default=0,
                    help="lanes per SpMV row; 0 = stream form")
    ap.add_argument("--config", default="c3",
                    choices=["c1", "c2", "c3", "c4a", "c4b", "c5", "factor"],
                    help="c3 (default) is the headline; the others print their own line")
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--nsys", type=int, default=48)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is design-only)", file=sys.stderr)
        return 2
    if args.config == "factor":
        return run_factor(args)
    import rchol_amd
    if rchol_amd.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    return {"c1": run_c1, "c2": run_c2, "c3": run_c3, "c4a": run_c4a, "c4b": run_c4b,
            "c5": run_c5}[args.config](args)


def run_c3(args):
    import rchol_amd
    from rchol_amd import factor as F

    s = args.side
    depth = int(np.log2(args.leaves)) + 1
    seeds = {"A": 20261202, "G": 1, "b": 7}
    log(f"generating {s}^3 variable-coefficient SDDM")
    A = F.varcoef_3d(s, s, s, 1.5, 1e-3, seeds["A"])
    n = A.shape[0]
    log(f"n={n} nnz={A.nnz}; nested dissection, {args.leaves} leaves")
    perm, blocks = F.nd_grid(s, s, s, depth)
    log(f"rchol factorization ({args.order} order inside ND blocks)")
    t0 = time.perf_counter()
    Gs, po, t_elim = F.factor(A, perm, blocks, args.order, seeds["G"])
    t_factor = time.perf_counter() - t0
    nnz_g = int(Gs[1].size)
    log(f"factor done in {t_factor:.1f}s (elimination {t_elim:.1f}s), nnz(G)={nnz_g}, "
        f"fill={2 * nnz_g / A.nnz:.3f}")
    P = po.astype(np.int64)
    Ap = F.permute(A, po)
    del A
    b = np.random.default_rng(seeds["b"]).uniform(-1.0, 1.0, n)
    bp = b[P]

    log("uploading (FP64 and FP32 factor contexts)")
    s64 = rchol_amd.Solver(Ap, Gs, rchol_amd.FP64, graph_iters=args.graph_iters,
                           spmv_lanes=args.spmv_lanes, trsv_lanes=args.trsv_lanes)
    s32 = rchol_amd.Solver(Ap, Gs, rchol_amd.FP32, graph_iters=args.graph_iters,
                           spmv_lanes=args.spmv_lanes, trsv_lanes=args.trsv_lanes)
    st64 = s64.stats()
    log(f"sweep levels fwd={st64['fwd_levels']} bwd={st64['bwd_levels']} "
        f"waves={st64['trsv_waves']}")
    setup = {name: dict(zip(("upload_a", "analysis", "schedule", "copies"), st["setup_ms"]))
             for name, st in (("fp64", st64), ("mixed", s32.stats()))}
    log("setup ms (mixed): " + ", ".join(f"{k} {v:.0f}" for k, v in setup["mixed"].items()))

    def run(solver, k):
        times, dev = [], []
        out = None
        for _ in range(k):
            t = time.perf_counter()
            out = solver.pcg(bp, args.tol, args.maxit)
            times.append(time.perf_counter() - t)
            dev.append(solver.stats()["last_solve_ms"] / 1e3)
        return out, times, dev

    log(f"warmup {args.warmup}")
    run(s32, args.warmup)
    run(s64, args.warmup)
    log(f"timed {args.steps} mixed solves")
    (x32, rr32, it32), t32, d32 = run(s32, args.steps)
    log(f"mixed: {it32} iterations, {np.mean(t32) * 1e3:.1f} ms/solve, relres {rr32:.2e}")
    (x64, rr64, it64), t64, d64 = run(s64, args.steps)
    log(f"fp64: {it64} iterations, {np.mean(t64) * 1e3:.1f} ms/solve, relres {rr64:.2e}")

    # per-kernel device time (HIP events on the engine's stream) over one
    # profiled solve of each variant
    prof = {}
    for name, solver, fp32 in (("mixed", s32, True), ("fp64", s64, False)):
        solver.set_option(rchol_amd.solver.OPT_PROFILE, 1)
        solver.pcg(bp, args.tol, args.maxit)
        st = solver.stats()
        solver.set_option(rchol_amd.solver.OPT_PROFILE, 0)
        iters = max(st["kernel_iters"], 1)
        avg_ms = [m / iters for m in st["kernel_ms"]]
        bts = kernel_bytes(n, Ap.nnz, nnz_g, fp32)
        prof[name] = {
            KERNEL_NAMES[i]: {"avg_ms": avg_ms[i], "bytes": bts[i],
                              "gbs": bts[i] / (avg_ms[i] * 1e-3) / 1e9 if avg_ms[i] > 0 else None}
            for i in range(4)}
        prof[name]["iters"] = st["kernel_iters"]
    log("per-kernel: " + json.dumps({k: {kk: round(vv["avg_ms"], 4) for kk, vv in v.items()
                                         if isinstance(vv, dict)} for k, v in prof.items()}))

    err = {}
    if not args.no_ref:
        log("FP64 reference solve to 1e-14")
        xref, rref, iref = s64.pcg(bp, 1e-14, 20000)
        nref = np.linalg.norm(xref)
        err = {"ref_tol": 1e-14, "ref_iters": iref, "ref_relres": rref,
               "mixed_err_vs_ref": float(np.linalg.norm(x32 - xref) / nref),
               "fp64_err_vs_ref": float(np.linalg.norm(x64 - xref) / nref)}
    # FP64-verified residual of the mixed solution
    r_true32 = float(np.linalg.norm(Ap @ x32 - bp) / np.linalg.norm(bp))

    # whole-iteration algorithmic bytes (solve-level GB/s)
    b32 = sum(kernel_bytes(n, Ap.nnz, nnz_g, True))
    b64 = sum(kernel_bytes(n, Ap.nnz, nnz_g, False))
    solve32 = float(np.mean(d32))
    solve64 = float(np.mean(d64))

    dom = max(KERNEL_NAMES, key=lambda k: prof["mixed"][k]["avg_ms"])
    dk = prof["mixed"][dom]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_mixed.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    copy = s32.copy_probe(1 << 30)
    log(f"plain copy: {copy['gbs']:.0f} GB/s")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["gbs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dk["gbs"] / HBM_PEAK_GBS, "traffic": traffic,
                "algorithmic_bytes": dk["bytes"], "avg_ms": dk["avg_ms"],
                "copy_gbs": copy["gbs"]}

    # the 32-right-hand-side batch (interleaved multivector) on the same
    # factor: per-RHS time against one FP64 single-RHS solve
    batch = None
    if args.batch_cols > 0:
        k = args.batch_cols
        B = np.random.default_rng(11).uniform(-1.0, 1.0, (n, k))
        log(f"batch of {k} right-hand sides (FP32 factor)")
        s32.pcg_batch(B, args.tol, args.maxit)
        tb = []
        for _ in range(max(1, args.batch_steps)):
            X, rrb, itb = s32.pcg_batch(B, args.tol, args.maxit)
            tb.append(s32.stats()["last_solve_ms"] / 1e3)
        worst = max(float(np.linalg.norm(Ap @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c]))
                    for c in range(k))
        del X
        tbm = float(np.mean(tb))
        batch = {"nrhs": k, "time_s": tbm, "per_rhs_s": tbm / k, "steps": len(tb),
                 "iters_min": int(itb.min()), "iters_max": int(itb.max()),
                 "max_true_relres": worst,
                 "speedup_per_rhs_vs_fp64_single": solve64 / (tbm / k)}
        log(f"batch: {tbm * 1e3:.0f} ms for {k} ({tbm / k * 1e3:.1f} ms per RHS, "
            f"{solve64 / (tbm / k):.2f}x one FP64 solve)")

    cpu = None
    parity = {}
    if not args.no_cpu:
        from oracle import oracle as O
        threads = min(16, os.cpu_count() or 1)
        prep = O.Prepared(Ap, Gs)
        # parity at full size: one FP64 preconditioner application and the
        # full FP64 solve against the oracle (the solve is also the CPU
        # baseline: sequential triangular sweeps as in the reference, OpenMP
        # SpMV and BLAS-1)
        O.set_threads(threads)
        z = s64.precond(bp)
        zo = prep.precond(bp)
        log(f"CPU baseline: full oracle FP64 PCG on {threads} threads")
        t = time.perf_counter()
        xo, ro, io = prep.pcg(bp, args.tol, args.maxit)
        tc = time.perf_counter() - t
        parity = {"oracle_iters": io, "fp64_iters": it64, "mixed_iters": it32,
                  "iters_within_1": abs(io - it64) <= 1,
                  "fp64_x_rel_diff_vs_oracle": float(np.linalg.norm(x64 - xo) /
                                                     np.linalg.norm(xo)),
                  "precond_fp64_rel_diff_vs_oracle": float(np.linalg.norm(z - zo) /
                                                           np.linalg.norm(zo)),
                  "oracle_relres": ro}
        log(f"CPU: {tc:.1f} s, {io} iterations (GPU FP64 {it64}); precond rel diff "
            f"{parity['precond_fp64_rel_diff_vs_oracle']:.1e}")
        if not parity["iters_within_1"] or parity["precond_fp64_rel_diff_vs_oracle"] > 1e-12:
            print(f"PARITY FAILURE at full size: {parity}", file=sys.stderr, flush=True)
        cpu = {"value": tc, "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"one full FP64 PCG solve of the same {s}^3 system to {args.tol:g} "
                          f"(oracle/pcg_oracle.c, the C restatement of pcg.cpp: sequential "
                          f"triangular sweeps, OpenMP SpMV and BLAS-1 on {threads} threads), "
                          f"{io} iterations"),
               "iters": io, "host": _host_info()}
        if not args.no_cpu_one_core:
            O.set_threads(1)
            t = time.perf_counter()
            _, _, io1 = prep.pcg(bp, args.tol, args.maxit)
            t1 = time.perf_counter() - t
            O.set_threads(threads)
            cpu["one_core"] = {"value": t1, "unit": "s", "cores": 1, "iters": io1}
            log(f"CPU, one core: {t1:.1f} s")

    # the drop-in path on the factor the reference's own order produces
    # (natural order inside the nested-dissection blocks): FP64, where the
    # engine plans chunked sweeps for its deep dependency chains
    static = None
    if not args.no_static:
        s32.close()
        s64.close()
        log("static order (the reference's order inside the ND blocks): factorization")
        A2 = F.varcoef_3d(s, s, s, 1.5, 1e-3, seeds["A"])
        t0 = time.perf_counter()
        Gst, pst, _ = F.factor(A2, perm, blocks, "static", seeds["G"])
        tfs = time.perf_counter() - t0
        Ast = F.permute(A2, pst)
        del A2
        bst = b[pst.astype(np.int64)]
        sst = rchol_amd.Solver(Ast, Gst, rchol_amd.FP64, graph_iters=args.graph_iters)
        stt = sst.stats()
        log(f"static: nnz(G)={Gst[1].size}, levels {stt['fwd_levels']}, chunked {stt['chunked']} "
            f"({stt['chunks'][0]} chunks), setup ms {[round(v) for v in stt['setup_ms']]}")
        sst.pcg(bst, args.tol, args.maxit)
        ts = []
        for _ in range(max(1, args.static_steps)):
            xs_, rrs, its = sst.pcg(bst, args.tol, args.maxit)
            ts.append(sst.stats()["last_solve_ms"] / 1e3)
        sst.set_option(rchol_amd.solver.OPT_PROFILE, 1)
        sst.pcg(bst, args.tol, args.maxit)
        stp = sst.stats()
        sst.close()
        kit = max(stp["kernel_iters"], 1)
        static = {"time_s": float(np.mean(ts)), "iters": its, "relres": rrs,
                  "relres_fp64_verified": float(np.linalg.norm(Ast @ xs_ - bst) /
                                                np.linalg.norm(bst)),
                  "nnz_g": int(Gst[1].size), "fill": 2 * Gst[1].size / Ast.nnz,
                  "levels": [stt["fwd_levels"], stt["bwd_levels"]], "chunked": stt["chunked"],
                  "chunks": stt["chunks"], "chunk_rounds": stt["chunk_rounds"],
                  "factor_s": tfs,
                  "setup_ms": dict(zip(("upload_a", "analysis", "schedule", "copies"),
                                       stt["setup_ms"])),
                  "kernel_ms_per_iter": {KERNEL_NAMES[i]: stp["kernel_ms"][i] / kit
                                         for i in range(4)}}
        log(f"static: {its} iterations, {static['time_s'] * 1e3:.0f} ms/solve (FP64)")
        del Gst, Ast

    res = {
        "metric": "PCG time-to-1e-10 & HBM GB/s, 256^3 var-coef Poisson; mixed vs FP64 GPU",
        "value": solve32,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": solve32 * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 (SpMV, dots, alpha/beta) + f32 (factor sweeps)",
        "data": "synthetic",
        "config": {
            "workload": (f"{s}^3 variable-coefficient 7-point Poisson (face conductivity "
                         f"exp(N(0,1.5^2)), +1e-3 I), ND {args.leaves} leaves, rchol factor "
                         f"({args.order} order in blocks), b ~ U[-1,1], PCG to {args.tol:g}"),
            "n": n, "nnz_a": int(Ap.nnz), "nnz_g": nnz_g, "fill": 2 * nnz_g / Ap.nnz,
            "leaves": args.leaves, "order": args.order, "seeds": seeds,
            "fwd_levels": st64["fwd_levels"], "bwd_levels": st64["bwd_levels"],
            "factor_s": t_factor, "parallelism": "single GPU",
        },
        "mixed": {"time_s": solve32, "host_time_s": float(np.mean(t32)), "iters": it32,
                  "relres": rr32, "relres_fp64_verified": r_true32,
                  "hbm_gbs": b32 * it32 / solve32 / 1e9, "kernels": prof["mixed"],
                  "setup_ms": setup["mixed"]},
        "fp64": {"time_s": solve64, "host_time_s": float(np.mean(t64)), "iters": it64,
                 "relres": rr64, "hbm_gbs": b64 * it64 / solve64 / 1e9,
                 "kernels": prof["fp64"], "setup_ms": setup["fp64"]},
        "speedup_mixed_vs_fp64": solve64 / solve32,
        "error": err,
        "parity": parity,
        "batch": batch,
        "static_order_fp64": static,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    _emit(res)
    s32.close()
    s64.close()
    return 0


def _host_info():
    """nproc, CPU model and OMP_NUM_THREADS of the host the baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"nproc": os.cpu_count(), "affinity": affinity, "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _emit(res):
    print(json.dumps(res), flush=True)


def _base(metric, value, unit, steps, warmup, hib, config, dtype):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": value * 1e3 if unit == "s" else None,
            "higher_is_better": hib, "scaling": "strong", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": config}


def _timed(fn, k):
    out, dev = None, []
    for _ in range(k):
        out = fn()
        dev.append(out[-1])
    return out, float(np.mean(dev))


def run_c1(args):
    """32^3 Laplacian, sequential rchol (natural order), FP64 GPU vs the oracle."""
    import rchol_amd
    from rchol_amd import factor as F
    from oracle import oracle as O
    side = 32 if args.side == 256 else args.side
    A = F.laplace_3d(side)
    n = A.shape[0]
    Gs, _, _ = F.factor(A, None, [n], "static", 1)
    b = np.random.default_rng(7).uniform(-1, 1, n)
    with rchol_amd.Solver(A, Gs, graph_iters=args.graph_iters) as s:
        for _ in range(args.warmup):
            s.pcg(b, 1e-10, 5000)
        def one():
            x, rr, it = s.pcg(b, 1e-10, 5000)
            return x, rr, it, s.stats()["last_solve_ms"] / 1e3
        (x, rr, it, _), t = _timed(one, args.steps)
        st = s.stats()
    t0 = time.perf_counter()
    xo, ro, io = O.pcg(A, b, 1e-10, 5000, Gs)
    tc = time.perf_counter() - t0
    res = _base("PCG time-to-1e-10, 32^3 Laplacian (sequential rchol)", t, "s", args.steps,
                args.warmup, False,
                {"workload": f"{side}^3 7-point Laplacian (laplace_3d), sequential rchol, "
                             "natural order, seed 1, b ~ U[-1,1] seed 7, FP64",
                 "n": n, "nnz_g": int(Gs[1].size), "fwd_levels": st["fwd_levels"]}, "f64")
    res.update({"iters": it, "oracle_iters": io, "relres": rr, "oracle_relres": ro,
                "x_rel_diff_vs_oracle": float(np.linalg.norm(x - xo) / np.linalg.norm(xo)),
                "cpu_baseline": {"value": tc, "unit": "s", "cores": 1, "kind": "port",
                                 "sample": "full oracle PCG solve (C restatement of pcg.cpp)"}})
    _emit(res)
    return 0


def _solve_pair(Ap, Gs, bp, args, tag):
    import rchol_amd
    out = {}
    for name, prec in (("fp64", rchol_amd.FP64), ("mixed", rchol_amd.FP32)):
        with rchol_amd.Solver(Ap, Gs, prec, graph_iters=args.graph_iters) as s:
            for _ in range(args.warmup):
                s.pcg(bp, args.tol, args.maxit)
            def one():
                x, rr, it = s.pcg(bp, args.tol, args.maxit)
                return x, rr, it, s.stats()["last_solve_ms"] / 1e3
            (x, rr, it, _), t = _timed(one, args.steps)
            st = s.stats()
            if name == "fp64":
                xref, _, iref = s.pcg(bp, 1e-14, 20000)
            out[name] = {"time_s": t, "iters": it, "relres": rr, "x": x,
                         "levels": (st["fwd_levels"], st["bwd_levels"])}
        log(f"{tag} {name}: {it} iterations, {t * 1e3:.1f} ms")
    for name in ("fp64", "mixed"):
        out[name]["err_vs_fp64_1e-14"] = float(np.linalg.norm(out[name].pop("x") - xref) /
                                               np.linalg.norm(xref))
    out["ref_iters_1e-14"] = iref
    return out


def run_c2(args):
    """128^3 Laplacian, ND 64 leaves: FP64 vs FP32 factor, both orderings."""
    from rchol_amd import factor as F
    side = 128 if args.side == 256 else args.side
    A = F.laplace_3d(side)
    n = A.shape[0]
    b = np.random.default_rng(7).uniform(-1, 1, n)
    perm, blocks = F.nd_grid(side, side, side, 7)
    res = None
    for order in ("mis", "static"):
        Gs, po, tf = F.factor(A, perm, blocks, order, 1)
        P = po.astype(np.int64)
        Ap = F.permute(A, po)
        r = _solve_pair(Ap, Gs, b[P], args, f"c2/{order}")
        r["fill"] = 2 * Gs[1].size / A.nnz
        if res is None:
            res = _base("PCG time-to-1e-10, 128^3 Laplacian ND 64 leaves; mixed vs FP64",
                        r["mixed"]["time_s"], "s", args.steps, args.warmup, False,
                        {"workload": f"{side}^3 laplace_3d, ND 64 leaves, rchol seed 1, "
                                     "b ~ U[-1,1] seed 7", "n": n}, "f64 + f32 factor")
        res[order] = r
    _emit(res)
    return 0


def run_c4a(args):
    """C3 matrix, a block of nrhs right-hand sides: per-RHS time vs FP64 single RHS."""
    import rchol_amd
    from rchol_amd import factor as F
    s = args.side
    depth = int(np.log2(args.leaves)) + 1
    A = F.varcoef_3d(s, s, s, 1.5, 1e-3, 20261202)
    n = A.shape[0]
    perm, blocks = F.nd_grid(s, s, s, depth)
    log("c4a: factor")
    Gs, po, _ = F.factor(A, perm, blocks, args.order, 1)
    Ap = F.permute(A, po)
    del A
    k = args.nrhs
    B = np.random.default_rng(11).uniform(-1, 1, (n, k))
    out = {}
    for name, prec in (("fp64", rchol_amd.FP64), ("mixed", rchol_amd.FP32)):
        with rchol_amd.Solver(Ap, Gs, prec, graph_iters=args.graph_iters) as sv:
            x1, rr1, it1 = sv.pcg(B[:, 0].copy(), args.tol, args.maxit)
            t1 = []
            for _ in range(args.steps):
                sv.pcg(B[:, 0].copy(), args.tol, args.maxit)
                t1.append(sv.stats()["last_solve_ms"] / 1e3)
            sv.pcg_batch(B, args.tol, args.maxit)
            tb = []
            for _ in range(args.steps):
                X, rr, it = sv.pcg_batch(B, args.tol, args.maxit)
                tb.append(sv.stats()["last_solve_ms"] / 1e3)
            true = [float(np.linalg.norm(Ap @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c]))
                    for c in range(k)]
            # per-kernel device time of the batched loop (HIP events)
            sv.set_option(rchol_amd.solver.OPT_PROFILE, 1)
            sv.pcg_batch(B, args.tol, args.maxit)
            stp = sv.stats()
            sv.set_option(rchol_amd.solver.OPT_PROFILE, 0)
            kms = {KERNEL_NAMES[i]: stp["kernel_ms"][i] / max(stp["kernel_iters"], 1)
                   for i in range(4)}
            out[name] = {"single_rhs_s": float(np.mean(t1)), "single_iters": it1,
                         "batch_s": float(np.mean(tb)), "per_rhs_s": float(np.mean(tb)) / k,
                         "iters_max": int(it.max()), "iters_min": int(it.min()),
                         "max_true_relres": max(true), "batch_kernel_ms_per_iter": kms}
            log(f"c4a {name}: single {out[name]['single_rhs_s']*1e3:.1f} ms, batch of {k} "
                f"{out[name]['batch_s']*1e3:.1f} ms; per iteration "
                + ", ".join(f"{kk} {vv:.2f} ms" for kk, vv in kms.items()))
    per_rhs = out["mixed"]["per_rhs_s"]
    res = _base("PCG per-RHS time-to-1e-10, 32-RHS batch on the 256^3 var-coef problem",
                per_rhs, "s", args.steps, 0, False,
                {"workload": f"{s}^3 var-coef Poisson, {k} RHS interleaved, FP32 factor batch "
                             "vs FP64 single RHS", "n": n, "nrhs": k}, "f64 + f32 factor")
    res.update(out)
    res["speedup_per_rhs_vs_fp64_single"] = out["fp64"]["single_rhs_s"] / per_rhs
    _emit(res)
    return 0


def run_c4b(args):
    """nsys uneven 2D 5-point SDDM systems in one launch per stage vs one by one."""
    import rchol_amd
    from rchol_amd import factor as F
    rng = np.random.default_rng(48)
    sides = np.exp(rng.uniform(np.log(64), np.log(1024), args.nsys)).astype(int)
    systems, bs = [], []
    t0 = time.perf_counter()
    for i, sd in enumerate(sides):
        A = F.varcoef_3d(int(sd), int(sd), 1, 1.0, 0.0, 1000 + i)
        perm, blocks = F.nd_grid(int(sd), int(sd), 1, 4)
        Gs, po, _ = F.factor(A, perm, blocks, "mis", i + 1)
        P = po.astype(np.int64)
        systems.append((A[P][:, P].tocsr(), Gs))
        bs.append(rng.uniform(-1, 1, A.shape[0]))
    log(f"c4b: {args.nsys} systems factored in {time.perf_counter() - t0:.1f}s, "
        f"n total {sum(a.shape[0] for a, _ in systems)}")
    Au, Gu, off = rchol_amd.stack_systems(systems)
    b = np.concatenate(bs)
    with rchol_amd.Solver(Au, Gu, rchol_amd.FP32, graph_iters=args.graph_iters) as sv:
        tune = sv.autotune(reps=3)
        sv.set_segments(off)
        sv.pcg_segments(b, args.tol, args.maxit)
        tb = []
        for _ in range(args.steps):
            x, rr, it = sv.pcg_segments(b, args.tol, args.maxit)
            tb.append(sv.stats()["last_solve_ms"] / 1e3)
    t_batch = float(np.mean(tb))
    t_seq = 0.0
    worst = 0.0
    for k, (A, Gs) in enumerate(systems):
        lo, hi = int(off[k]), int(off[k + 1])
        worst = max(worst, float(np.linalg.norm(A @ x[lo:hi] - b[lo:hi]) /
                                 np.linalg.norm(b[lo:hi])))
        with rchol_amd.Solver(A, Gs, rchol_amd.FP32, graph_iters=args.graph_iters) as sv:
            sv.pcg(bs[k], args.tol, args.maxit)
            sv.pcg(bs[k], args.tol, args.maxit)
            t_seq += sv.stats()["last_solve_ms"] / 1e3
    res = _base("uneven batch of 2D 5-point SDDM systems to 1e-10: one launch per stage",
                t_batch, "s", args.steps, 1, False,
                {"workload": f"{args.nsys} 2D 5-point SDDM, sides log-uniform 64..1024, "
                             "log-normal weights sigma=1, FP32 factor, ND 8 leaves + MIS",
                 "sides": sides.tolist(), "n_total": int(off[-1])}, "f64 + f32 factor")
    res.update({"sequential_s": t_seq, "speedup_vs_sequential": t_seq / t_batch,
                "iters_max": int(it.max()), "iters_min": int(it.min()),
                "max_true_relres": worst, "autotune": tune})
    _emit(res)
    return 0


def run_factor(args):
    """Host only: this package's factorization (librchol_factor.so) against the
    reference's own (c++/rchol_lap, built from its sources into oracle/_ref),
    on the same matrix, nested dissection and seed.  --side 128 --leaves 64 is
    the 7-point Laplacian of the 128^3 workload; other sides use the
    variable-coefficient problem of the headline.  Each side runs on the
    threads it spawns (one per leaf, as rchol_lap.cpp:186-263) plus this
    package's host-parallel loops; the reference pins its leaf threads to
    cores 0..leaves-1 (rchol_lap.cpp:100-114, :190-197).  The static-order
    factor is checked bit-identical to the reference's."""
    import resource
    from rchol_amd import factor as F
    from oracle import oracle as O
    s, leaves, seed = args.side, args.leaves, 1
    depth = int(np.log2(leaves)) + 1
    if s == 128:
        A, what = F.laplace_3d(s), "7-point Laplacian"
    else:
        A, what = F.varcoef_3d(s, s, s, 1.5, 1e-3, 20261202), "variable-coefficient Poisson"
    perm, blocks = F.nd_grid(s, s, s, depth)
    rows = {}
    for order in ("static", "mis"):
        t = time.perf_counter()
        Gs, po, te = F.factor(A, perm, blocks, order, seed)
        rows[order] = {"wall_s": time.perf_counter() - t, "elimination_s": te,
                       "nnz_g": int(Gs[0][-1])}
        if order == "static":
            G_static = Gs
        else:
            del Gs
        log(f"factor {order}: {rows[order]}")
    ref = None
    if O.ref_available() and not args.no_ref:
        t = time.perf_counter()
        Gr = O.ref_rchol_nd(A, perm, blocks, seed)
        ref = {"wall_s": time.perf_counter() - t, "nnz_g": int(Gr[0][-1]),
               "bit_identical_to_static": all(np.array_equal(a, b) for a, b in zip(Gr, G_static))}
        log(f"reference: {ref}")
    res = _base("factor_seconds", rows["static"]["wall_s"], "s", 1, 0, False,
                {"workload": f"{s}^3 {what}, ND {leaves} leaves, seed {seed}",
                 "n": int(A.shape[0]), "nnz_a": int(A.nnz), "leaves": leaves}, "f64")
    res.update({"this_static": rows["static"], "this_mis": rows["mis"], "reference": ref,
                "speedup_static_vs_reference": ref["wall_s"] / rows["static"]["wall_s"] if ref else None,
                "max_rss_gb": resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 2**20,
                "host": _host_info()})
    _emit(res)
    return 0


def run_c5(args):
    """100^3 weighted graph Laplacian: differentiable solve, mixed forward +
    adjoint.  Two timings: a training step whose weights change every step
    (assembly, factorization and upload on the host, then the device forward
    and adjoint: what an optimiser sees), and a step that reuses the factor of
    unchanged weights.  The gradient is checked against an FP64 adjoint at
    full size."""
    import torch
    import rchol_amd
    from rchol_amd import autograd as AG
    side = 100 if args.side == 256 else args.side
    dev = torch.device("cuda", 0)
    n = side ** 3
    b0 = np.random.default_rng(5).uniform(-1, 1, n)
    xs = np.random.default_rng(2).uniform(-1, 1, n)
    bt = torch.tensor(b0, device=dev)
    xt = torch.tensor(xs, device=dev)
    grads, cached, fresh, split = {}, {}, {}, {}
    for name, prec in (("mixed", rchol_amd.FP32), ("fp64", rchol_amd.FP64)):
        prob = AG.GridLaplacian(side, side, side, shift=1e-2, leaves=16, precision=prec,
                                tol=1e-10 if name == "mixed" else 1e-13)
        theta0 = np.random.default_rng(1).standard_normal(prob.m)
        theta = torch.tensor(theta0, device=dev, requires_grad=True)

        def step():
            theta.grad = None
            w = torch.nn.functional.softplus(theta)
            x = AG.solve(w, bt, prob)
            ((x - xt) ** 2).sum().backward()

        # untimed: the first step factors, uploads and captures the graph;
        # the next ones run the cached-factor path once (its first use of
        # torch kernels loads their code)
        for _ in range(max(args.warmup, 1) + 1):
            step()
        torch.cuda.synchronize()
        grads[name] = theta.grad.cpu().numpy()
        t = time.perf_counter()
        for _ in range(max(args.steps, 1)):
            step()
        torch.cuda.synchronize()
        cached[name] = (time.perf_counter() - t) / max(args.steps, 1)
        # weights that change every step (an optimiser update in between)
        lr = 1e-3
        t = time.perf_counter()
        tf = 0.0
        for _ in range(max(args.steps, 1)):
            with torch.no_grad():
                theta -= lr * theta.grad
            t1 = time.perf_counter()
            prob.solver_for(torch.nn.functional.softplus(theta).detach())
            tf += time.perf_counter() - t1
            step()
        torch.cuda.synchronize()
        fresh[name] = (time.perf_counter() - t) / max(args.steps, 1)
        split[name] = {"factor_and_upload_s": tf / max(args.steps, 1),
                       "device_forward_adjoint_s": fresh[name] - tf / max(args.steps, 1)}
        log(f"c5 {name}: forward {prob.last['forward']}, adjoint {prob.last['adjoint']}, "
            f"cached-factor step {cached[name]:.3f}s, new-weights step {fresh[name]:.3f}s")
    g, gr = grads["mixed"], grads["fp64"]
    res = _base("differentiable solve: training step (forward + adjoint) time, 100^3 graph "
                "Laplacian", fresh["mixed"], "s", max(args.steps, 1), max(args.warmup, 1) + 1,
                False,
                {"workload": f"{side}^3 6-neighbour graph Laplacian, w = softplus(theta), "
                             "+1e-2 I, loss ||x - x*||^2, mixed forward/adjoint PCG; theta "
                             "changes every step (host assembly + rchol + upload each step)",
                 "n": n, "edges": int(g.size)}, "f64 + f32 factor")
    res.update({"new_weights_step_s": fresh, "cached_factor_step_s": cached,
                "new_weights_split": split,
                "grad_rel_diff_vs_fp64_adjoint": float(np.linalg.norm(g - gr) /
                                                       np.linalg.norm(gr))})
    _emit(res)
    return 0


if __name__ == "__main__":
    sys.exit(main())
