"""The host analysis of a stored factor (rchol_amd_analyze_G, the code
upload_G runs) against an independent restatement: the dependency level of
a row is the length of the longest chain of rows it waits for, and slots
are the rows ordered by (level, row).  The restatement relaxes all entries
at once until nothing changes (a Bellman-Ford longest path), so it shares
no code or order of work with the threaded piece scheduler it checks.
Shapes cover one row, no off-diagonal entries, a star (one row read by all,
reading all), a chain, rows alternating between reading and not reading
(more nested-dissection block starts than the scheduler keeps), and
factors large enough to split over host threads.  The per-slot proxies of
the sweep's waits are restated from the same levels, without the entry
order they are read from."""
import numpy as np
import pytest

import rchol_amd
from rchol_amd import factor as F


def tight_levels(cp, ri, n):
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(np.asarray(cp, dtype=np.int64)))
    rows = np.asarray(ri, dtype=np.int64)
    off = rows != cols
    out = {}
    for key, dst, src in (("fwd", rows[off], cols[off]), ("bwd", cols[off], rows[off])):
        lev = np.zeros(n, dtype=np.int64)
        while True:
            new = np.zeros(n, dtype=np.int64)
            np.maximum.at(new, dst, lev[src] + 1)
            if np.array_equal(new, lev):
                break
            lev = new
        out[key] = (lev, dst, src)
    return out


def row_proxies(lev, dst, src):
    """Per row of level l >= 2: c = its dependencies of level l - 1 (a row
    stores them last), kept when its deepest other dependency is at least
    level l - 3 (the row's entry just before the trailing ones), else 0."""
    n = lev.size
    dl, rl = lev[src], lev[dst]
    c = np.bincount(dst[dl == rl - 1], minlength=n)
    below = np.full(n, -1, dtype=np.int64)
    m = dl < rl - 1
    np.maximum.at(below, dst[m], dl[m])
    ok = (lev >= 2) & (below >= 0) & (below >= lev - 3) & (c <= 0xFFFF)
    return np.where(ok, c, 0)


def check(Gs):
    cp, ri, gv = (np.asarray(a) for a in Gs)
    n = cp.size - 1
    sch = rchol_amd.analyze((cp.astype(np.uint64), ri.astype(np.uint64), gv.astype(np.float64)), 0)
    want = tight_levels(cp, ri, n)
    for key in ("fwd", "bwd"):
        lev, dst, src = want[key]
        s = sch[key]
        assert s["nlevels"] == int(lev.max()) + 1
        np.testing.assert_array_equal(s["level_ptr"],
                                      np.concatenate([[0], np.cumsum(np.bincount(lev))]))
        np.testing.assert_array_equal(s["slot_row"], np.lexsort((np.arange(n), lev)))
        # proxies: per slot, those of rows in levels of fewer than 2^18 rows
        width = np.bincount(lev)[lev]
        px = np.where(width < (1 << 18), row_proxies(lev, dst, src), 0)
        np.testing.assert_array_equal(s["proxy"], px[s["slot_row"]])


def lower(n, rows, cols, seed=0):
    """Stored arrays of a lower factor with the given off-diagonal pattern."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    assert np.all(rows > cols)
    r = np.concatenate([np.arange(n), rows])
    c = np.concatenate([np.arange(n), cols])
    o = np.lexsort((r, c))
    r, c = r[o], c[o]
    v = np.random.default_rng(seed).uniform(0.5, 1.5, r.size)
    cp = np.concatenate([[0], np.cumsum(np.bincount(c, minlength=n))])
    return cp, r, v


def test_one_row():
    check((np.array([0, 1]), np.array([0]), np.array([2.0])))


def test_no_off_diagonal_entries():
    n = 50000
    check(lower(n, [], []))


def test_star():
    n = 40000
    check(lower(n, np.arange(1, n), np.zeros(n - 1, dtype=np.int64)))


def test_chain():
    n = 30000
    check(lower(n, np.arange(1, n), np.arange(n - 1)))


def test_alternating_rows_exceed_the_piece_cap():
    # odd rows read the row before; even rows read nothing: a block start
    # every second row, far more than the scheduler's piece limit
    n = 200000
    odd = np.arange(1, n, 2)
    check(lower(n, odd, odd - 1))


def test_random_lower_pattern():
    rng = np.random.default_rng(5)
    n = 60000
    i = rng.integers(1, n, 300000)
    j = (i * rng.uniform(0, 1, i.size)).astype(np.int64)
    key = np.unique(i * n + j)
    check(lower(n, key // n, key % n))


@pytest.mark.parametrize("order", ["static", "mis"])
def test_dissected_factors_split_over_threads(order):
    A = F.varcoef_3d(32, 32, 32, 1.5, 1e-3, 4)
    perm, blocks = F.nd_grid(32, 32, 32, 4)
    Gs, _, _ = F.factor(A, perm, blocks, order, 6)
    check(Gs)
