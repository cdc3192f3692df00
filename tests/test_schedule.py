"""Host-side level schedule of the triangular sweeps (CPU, no device).

The persistent sweep kernel is deadlock-free only if every dependency of a
row lies in a strictly lower level of its sweep and the schedule covers every
row exactly once; the split head/tail phases need levels of at most the
threshold width.  Checked on the golden fixtures and on the package's
factors through rchol_amd_analyze_G (the code upload_G runs)."""
import glob
import os

import numpy as np
import pytest

import rchol_amd
from rchol_amd import factor as F

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz")))


def factors():
    out = []
    for name in FIXTURES:
        d = np.load(os.path.join(GOLD, name + ".npz"))
        out.append((name, (d["g_colptr"], d["g_rowidx"], d["g_val"])))
    A = F.varcoef_3d(14, 13, 12, 1.5, 1e-3, 3)
    perm, blocks = F.nd_grid(14, 13, 12, 4)
    for order in ("static", "mis"):
        Gs, _, _ = F.factor(A, perm, blocks, order, 2)
        out.append((f"var14_{order}", Gs))
    return out


CASES = factors()


@pytest.mark.parametrize("name,Gs", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("width", [0, 8, 64])
def test_schedule_respects_dependencies(name, Gs, width):
    cp, ri, _ = (np.asarray(a) for a in Gs)
    n = cp.size - 1
    sch = rchol_amd.analyze(Gs, width)
    cols = np.repeat(np.arange(n), np.diff(cp.astype(np.int64)))
    rows = ri.astype(np.int64)
    off = rows != cols
    for k, key in enumerate(("fwd", "bwd")):
        s = sch[key]
        lp, sr = s["level_ptr"], s["slot_row"]
        assert lp[0] == 0 and lp[-1] == n and np.all(np.diff(lp) > 0)
        assert np.array_equal(np.sort(sr), np.arange(n))
        lev = np.empty(n, dtype=np.int64)
        lev[sr] = np.repeat(np.arange(s["nlevels"]), np.diff(lp))
        if key == "fwd":   # row i waits for column j < i
            assert np.all(lev[rows[off]] > lev[cols[off]])
        else:              # row j of G^T waits for i > j
            assert np.all(lev[cols[off]] > lev[rows[off]])
        widths = np.diff(lp)
        h, t = s["head_end"], s["tail_begin"]
        assert 0 <= h <= t <= s["nlevels"]
        if width == 0:
            assert h == 0 and t == s["nlevels"]
        else:
            assert np.all(widths[:h] <= width) and np.all(widths[t:] <= width)
            assert h == s["nlevels"] or widths[h] > width
        assert s["mid_tasks"] == int(np.sum((widths[h:t] + 63) // 64))


def test_schedule_rejects_invalid_factors():
    with pytest.raises(RuntimeError, match="error -1"):
        rchol_amd.analyze((np.array([0, 1, 3]), np.array([0, 0, 1]), np.ones(3)))  # upper entry
    cp = np.array([0, 2, 3])
    with pytest.raises(RuntimeError, match="error -5"):
        rchol_amd.analyze((cp, np.array([0, 1, 1]), np.array([0.0, 1.0, 1.0])))  # zero pivot
    with pytest.raises(RuntimeError, match="error -1"):
        rchol_amd.analyze((cp, np.array([0, 0, 1]), np.ones(3)))  # two diagonals


def test_threaded_analysis_matches_one_thread(monkeypatch):
    """At 2^18 rows the analysis splits columns and slots over host threads;
    its schedule must not depend on the thread count, and a malformed
    factor must be reported at its lowest bad column, as one thread would."""
    A = F.varcoef_3d(64, 64, 64, 1.5, 1e-3, 9)
    perm, blocks = F.nd_grid(64, 64, 64, 5)
    Gs, _, _ = F.factor(A, perm, blocks, "mis", 3)
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    one = rchol_amd.analyze(Gs, 0)
    monkeypatch.setenv("OMP_NUM_THREADS", "8")
    many = rchol_amd.analyze(Gs, 0)
    for key in ("fwd", "bwd"):
        assert one[key]["nlevels"] == many[key]["nlevels"]
        assert np.array_equal(one[key]["level_ptr"], many[key]["level_ptr"])
        assert np.array_equal(one[key]["slot_row"], many[key]["slot_row"])
    cp, ri, gv = (np.array(a) for a in Gs)
    n = cp.size - 1
    bad = gv.copy()
    for c in (n - 5, n // 3, n // 2):   # zero pivots; the lowest must be named
        bad[int(cp[c])] = 0.0
    with pytest.raises(RuntimeError, match=rf"G\({n // 3},{n // 3}\) is zero"):
        rchol_amd.analyze((cp, ri, bad), 0)
    cp2 = cp.copy()
    cp2[n // 4 + 1] = cp2[n // 4] - 1        # a decreasing column pointer
    with pytest.raises(RuntimeError, match=rf"not monotone at {n // 4}"):
        rchol_amd.analyze((cp2, ri, gv), 0)
