"""The chunked sweeps' host plan (rchol_amd_chunk_apply_host: the plan run on
the CPU in the device kernel's order of operations, with its invariants
checked) against the oracle's preconditioner application, on the golden
fixtures and on deep factors (natural order, the reference's order inside
nested-dissection blocks), for chunk sizes that force many cuts."""
import glob
import os

import numpy as np
import pytest

from rchol_amd import factor as F
from rchol_amd.solver import chunk_apply_host
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz")))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("rows", [1, 7, 64, 300, 65535])
def test_plan_matches_oracle_on_fixtures(name, rows):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    Gs = (d["g_colptr"], d["g_rowidx"], d["g_val"])
    z, cnt, nrd, est = chunk_apply_host(Gs, d["b"], rows)
    assert rel(z, d["z"]) <= 1e-13
    n = d["g_colptr"].size - 1
    assert 1 <= cnt[0] <= n


@pytest.mark.parametrize("case", ["natural16", "static_nd20", "mis_nd20", "varcoef_static12"])
@pytest.mark.parametrize("rows", [256, 1000, 65535])
def test_plan_matches_oracle_on_deep_factors(case, rows):
    if case == "natural16":
        A = F.laplace_3d(16)
        Gs, _, _ = F.factor(A, None, [A.shape[0]], "static", 3)
    elif case == "varcoef_static12":
        A = F.varcoef_3d(12, 12, 12, 1.5, 1e-3, 4)
        perm, blocks = F.nd_grid(12, 12, 12, 3)
        Gs, _, _ = F.factor(A, perm, blocks, "static", 2)
    else:
        A = F.laplace_3d(20)
        perm, blocks = F.nd_grid(20, 20, 20, 4)
        Gs, _, _ = F.factor(A, perm, blocks, case.split("_")[0], 5)
    r = np.random.default_rng(1).uniform(-1, 1, A.shape[0])
    z, cnt, nrd, est = chunk_apply_host(Gs, r, rows)
    assert rel(z, O.precond(Gs, r)) <= 1e-12
    if rows == 256 and case == "static_nd20":
        # eight leaves at least: the cuts found the nested-dissection blocks
        assert cnt[0] >= 8


@pytest.mark.parametrize("rows", [4096, 65535])
def test_plan_with_external_parts_matches_oracle(rows):
    """Separators of a 40^3 factor in the reference's order read enough
    external entries that their sums go to parts (separate work items)."""
    A = F.laplace_3d(40)
    perm, blocks = F.nd_grid(40, 40, 40, 4)
    Gs, _, _ = F.factor(A, perm, blocks, "static", 7)
    r = np.random.default_rng(2).uniform(-1, 1, A.shape[0])
    z, cnt, nrd, est = chunk_apply_host(Gs, r, rows)
    assert cnt[1] > 0 and cnt[2] > 0
    assert rel(z, O.precond(Gs, r)) <= 1e-12
