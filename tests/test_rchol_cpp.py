"""The C++ factorization entry points of include/rchol_amd_rchol.hpp:
rchol(A, G) (c++/rchol/rchol.hpp:6) and rchol(A, G, P, threads[, part])
(c++/rchol/rchol_parallel.hpp:6), compiled into tests/cpp/rchol_cpp_check.cpp.

Pinned against the reference's own factorization (oracle/_ref, built from
c++/rchol_lap sources) on the same ordering and seed: G is bit-identical.
The block boundaries `part` follow rchol_parallel.cpp:63-69 (0, prefix sums
of the post-order block sizes, last + 1); argument errors carry the
reference's exception texts (rchol_parallel.cpp:39-43,
find_separator.cpp:123-126).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from rchol_amd import factor as F
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIDE = 10


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    lib = os.path.join(ROOT, "rchol_amd", "librchol_factor.so")
    assert os.path.exists(lib), "librchol_factor.so is built by make -C rchol_amd/csrc factor"
    d = tmp_path_factory.mktemp("rcc")
    exe = d / "rchol_cpp_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-o",
                    str(exe), os.path.join(ROOT, "tests", "cpp", "rchol_cpp_check.cpp"),
                    "-L" + os.path.dirname(lib), "-lrchol_factor",
                    "-Wl,-rpath," + os.path.dirname(lib), "-pthread"], check=True)
    out = subprocess.run([str(exe), str(SIDE), str(d)], capture_output=True, text=True,
                         timeout=120, check=True).stdout

    def load(tag):
        cp = np.fromfile(d / f"{tag}_cp.bin", dtype=np.uint64)
        ri = np.fromfile(d / f"{tag}_ri.bin", dtype=np.uint64)
        gv = np.fromfile(d / f"{tag}_gv.bin", dtype=np.float64)
        p = np.fromfile(d / f"{tag}_p.bin", dtype=np.uint64)
        part = np.fromfile(d / f"{tag}_part.bin", dtype=np.uint64)
        return (cp, ri, gv), p, part
    return out, load


def part_of(blocks):
    """result_idx of rchol_parallel.cpp:63-69 restated."""
    part = [0]
    for b in blocks:
        part.append(part[-1] + int(b))
    part[-1] += 1
    return np.array(part, dtype=np.uint64)


def same_factor(G, H):
    for a, b in zip(G, H):
        np.testing.assert_array_equal(a, b)


def test_sequential_matches_reference(run):
    out, load = run
    A = F.laplace_3d(SIDE)
    n = A.shape[0]
    G, P, part = load("seq5")
    np.testing.assert_array_equal(P, np.arange(n))
    np.testing.assert_array_equal(part, [0, n + 1])
    if O.ref_available():
        same_factor(G, O.ref_rchol_seq(A, 5))
    same_factor(G, F.factor(A, None, [n], "static", 5)[0])


def test_dissected_matches_reference_and_returns_part(run):
    out, load = run
    A = F.laplace_3d(SIDE)
    G, P, part = load("graph9")
    perm, blocks = F.nd_graph(A, 3)
    np.testing.assert_array_equal(P, perm)
    np.testing.assert_array_equal(part, part_of(blocks))
    assert part.size == 8 and int(part[-1]) == A.shape[0] + 1
    if O.ref_available():
        same_factor(G, O.ref_rchol_nd(A, perm, blocks, 9))
    same_factor(G, F.factor(A, perm, blocks, "static", 9)[0])


def test_unseeded_parallel_appends_and_preconditions(run):
    out, load = run
    A = F.laplace_3d(SIDE)
    n = A.shape[0]
    G, P, part = load("par8")
    perm, blocks = F.nd_graph(A, 4)
    np.testing.assert_array_equal(P[:2], [7, 7])   # appended, as rchol_parallel.cpp:62
    np.testing.assert_array_equal(P[2:], perm)
    np.testing.assert_array_equal(part, part_of(blocks))
    cp, ri, gv = G
    assert cp.size == n + 1 and np.all(np.diff(cp) >= 1)
    col = np.repeat(np.arange(n, dtype=np.uint64), np.diff(cp).astype(np.int64))
    assert np.all(ri >= col)                         # lower triangle
    Ap = F.permute(A, P[2:])
    b = np.random.default_rng(3).uniform(-1, 1, n)
    x, relres, itr = O.pcg(Ap, b, 1e-8, 200, G)
    assert relres <= 1e-8 and itr < 60
    # rchol(A, G, P, 2) and rchol(A, G): sizes of G, P and nnz(G)
    word, n2, p2, n1, nnz1 = out.splitlines()[0].split()
    assert word == "sizes" and int(n2) == int(p2) == int(n1) == n and int(nnz1) > n


def test_argument_errors_carry_reference_texts(run):
    out, _ = run
    lines = dict(l.split(" ", 1) for l in out.splitlines()[1:])
    assert lines["0"] == "thread number should be positive"
    assert lines["-4"] == "thread number should be positive"
    assert lines["3"] == "thread number should be a power of 2"
    assert lines["6"] == "thread number should be a power of 2"
    assert lines[str(1 << 20)] == "too many threads requested"
