"""Host-side checks of libmfdag through its C-ABI (no GPU needed): the library
loads and exports every entry point include/mfdag.h declares, and the host
analysis gives exactly the integer structure of the oracle's restatement of
the reference analysis (supernode partition, row lists, parents, perm,
ccs_perm, supernodal_index, local_index)."""
import os
import re

import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from oracle import analysis as A
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    L = mfdag.lib()
    hdr = open(os.path.join(ROOT, "include", "mfdag.h")).read()
    names = set(re.findall(r"\b(mfd_[a-z_]+)\s*\(", hdr))
    assert len(names) >= 20
    for name in sorted(names):
        assert hasattr(L, name), name


def test_error_strings():
    L = mfdag.lib()
    assert L.mfd_error_string(0) == b"ok"
    assert b"positive definite" in L.mfd_error_string(7)


def host_solver(nb=64):
    return mfdag.Solver(nb=nb, host_only=True)


def jag(lengths, vals):
    out, k = [], 0
    for n in lengths:
        out.append(list(vals[k:k + n]))
        k += n
    return out


def as_ccs(ccs_col, ccs_row):
    return [list(ccs_row[ccs_col[j] - 1:ccs_col[j + 1] - 1]) for j in range(len(ccs_col) - 1)]


def compare_with_oracle(ccs_col, ccs_row, max_zero, perm=None):
    s = host_solver()
    s.analyze(ccs_col, ccs_row, max_zero, perm)
    got = s.structure()
    exp = A.hppllt_analyze(as_ccs(ccs_col, ccs_row), max_zero, None if perm is None else list(perm))
    assert got["supernode_size"].tolist() == exp["node_sizes"]
    assert got["work_size"].tolist() == exp["work_sizes"]
    assert got["parent"].tolist() == exp["parent"]
    assert got["perm"].tolist() == exp["perm"]
    assert got["ccs_perm"].tolist() == exp["ccs_perm"]
    assert got["supernodal_index"].tolist() == sum(exp["supernodal_index"], [])
    # local_index of the root is never written by the reference
    li = sum(exp["local_index"], [])
    assert got["local_index"].tolist() == li
    assert s.cost() == A.cost(exp["node_sizes"], exp["work_sizes"])
    assert s.nonzero() == A.nonzero(exp["node_sizes"], exp["work_sizes"])
    return s, got


def test_9x9_fixture(kats):
    k = kats["analysis_9x9"]
    cc = np.concatenate([[1], 1 + np.cumsum(k["ccs_num_row"])])
    for mz in (0, 1, 2, 5):
        compare_with_oracle(cc, k["ccs_row"], mz)


@pytest.mark.parametrize("seed", range(8))
def test_random_patterns(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(10, 120))
    m = rng.random((n, n)) < rng.uniform(0.02, 0.15)
    m = np.tril(m | m.T, -1)
    m[np.arange(1, n), np.arange(n - 1)] = True
    rows, cols = np.nonzero(m)
    r = np.concatenate([np.arange(n), rows])
    c = np.concatenate([np.arange(n), cols])
    cc, rr, _ = P.lower_ccs(n, r, c, np.ones(len(r)))
    perm = rng.permutation(n) + 1 if seed % 2 else None
    compare_with_oracle(cc, rr, int(rng.choice([0, 1, 4, 16, 64])), perm)


@pytest.mark.parametrize("shape,mz", [((12, 12), 0), ((20, 17), 8), ((6, 6, 6), 0), ((7, 5, 6), 32)])
def test_grids_with_nested_dissection(shape, mz):
    p = P.laplacian_2d(*shape) if len(shape) == 2 else P.weighted_laplacian_3d(*shape)
    compare_with_oracle(p.ccs_col, p.ccs_row, mz, p.perm)


def test_cost_kat_through_set_structure(kats):
    case = kats["cost"]["cases"][0]
    s = host_solver(nb=3)
    c, w = case["supernode_size"], case["work_size"]
    n = sum(c)
    colptr = np.arange(n + 1, dtype=np.int64)
    snidx = np.concatenate([np.arange(1, ci + 1) for ci in c])
    s.set_structure(c, w, [2, 0], [2, 3, 7, 8], colptr, snidx)
    assert s.cost() == case["cost"]
    assert s.nonzero() == case["nonzero"]


@pytest.mark.parametrize("nb", [1, 2, 3, 5, 8, 64])
def test_reference_array_sizes(nb):
    """hppllt_get_array_size(1..3) equals the oracle's array sizes."""
    p = P.weighted_laplacian_3d(5, 6, 4)
    s = host_solver(nb)
    s.analyze(p.ccs_col, p.ccs_row, 4, p.perm)
    st = s.structure()
    f = O.OracleFactor(st["supernode_size"], st["work_size"], st["parent"], nb, st["local_index"])
    for which in (1, 2, 3):
        assert s.array_size(which) == O.lib().orc_array_size(f.h, which)


def test_invalid_inputs():
    s = host_solver()
    with pytest.raises(mfdag.MfdError):
        s.analyze([1, 2, 3], [2, 2], 0)          # diagonal not first in column 1
    with pytest.raises(mfdag.MfdError) as e:
        s.analyze([1, 2, 3], [1, 2], 0)          # two disconnected columns
    assert e.value.code == -5
    with pytest.raises(mfdag.MfdError):
        s.analyze([1, 3, 4], [1, 2, 2], 0, perm=[1, 1])  # not a permutation
    if not s.device:
        with pytest.raises(mfdag.MfdError):
            s.analyze([1, 3, 4], [1, 2, 2], 0)
            s.factorize(np.array([2.0, -1.0, 2.0]))
