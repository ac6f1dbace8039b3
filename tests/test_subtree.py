"""Subtree partition of the assembly tree (mfd_tree_pruning; the reference's
tree_pruning, tree_pruning_m.f90:13-74), host only.

Pinned by the reference's known answers (tree_pruning_test.f90:24-33 on the
7-supernode fixture of tree_maker_m / sparse_matrix_maker_m, held in
tests/golden/reference_kats.json), which fix the list order (children added
at the head), the choice of the first heaviest subtree, the stopping ratio
and the fallback to ip - 1 workers; then the C-ABI against the oracle's
restatement on analysed problems, and what the partition must satisfy:
disjoint subtrees, at least p bins, the lightest bin above tol of the
heaviest (the reference stops below it only when the tree runs out of
subtrees, which the known answers cover)."""
import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from oracle import analysis as A

KAT_CASES = [(0.0, 1, [7]), (0.0, 2, [6, 2]), (0.71, 2, [5, 3, 2]), (0.73, 2, [5, 3, 2])]


def fixture(kats):
    k = kats["analysis_9x9"]
    tc, off = [], 0
    for nc in k["supernodal_tree_num_child"]:
        tc.append(k["supernodal_tree_child_val"][off:off + nc])
        off += nc
    parent = [0] * len(tc)
    for p, chs in enumerate(tc):
        for ch in chs:
            parent[ch - 1] = p + 1
    return k["supernodal_cc"], k["supernodal_node_size"], tc, parent


@pytest.mark.parametrize("tol,p,expected", KAT_CASES)
def test_reference_known_answers(kats, tol, p, expected):
    cc, ns, tc, parent = fixture(kats)
    assert A.tree_pruning(cc, ns, tc, tol, p) == expected
    sub, bins = mfdag.tree_pruning(ns, cc, parent, tol, p)
    assert sub.tolist() == expected


def tree_of(solver):
    st = solver.structure()
    c, w, parent = st["supernode_size"], st["work_size"], st["parent"]
    tc = [[] for _ in c]
    for s, pa in enumerate(parent):
        if pa > 0:
            tc[pa - 1].append(s + 1)
    roots = [s + 1 for s, pa in enumerate(parent) if pa == 0]
    return c, w, parent, tc, roots


def analysed(name):
    if name == "lap2d_30":
        p = P.laplacian_2d(30)
    elif name == "wlap3d_10":
        p = P.weighted_laplacian_3d(10, seed=1)
    else:
        p = P.stencil27_3d(9, seed=2)
    s = mfdag.Solver(nb=16, host_only=True)
    s.analyze(p.ccs_col, p.ccs_row, 8, p.perm)
    return s


def check_partition(c, w, parent, tc, sub, bins, tol, p, oracle):
    assert sub.tolist() == oracle
    # disjoint: no listed root lies in another's subtree
    listed = set(sub.tolist())
    for r in listed:
        a = parent[r - 1]
        while a > 0:
            assert a not in listed
            a = parent[a - 1]
    nbins = int(bins.max()) + 1
    assert nbins >= min(p, len(sub))
    wt = A.generate_tree_weights(A.count_op_count(w, c), tc)
    load = np.zeros(nbins)
    for r, b in zip(sub, bins):
        load[b] += wt[r - 1]
    assert load.min() > 0
    return load.min() / load.max()


@pytest.mark.parametrize("name", ["lap2d_30", "wlap3d_10", "st27_9"])
@pytest.mark.parametrize("tol,p", [(0.0, 1), (0.0, 4), (0.5, 2), (0.8, 4), (0.9, 8)])
def test_against_oracle(name, tol, p):
    s = analysed(name)
    c, w, parent, tc, roots = tree_of(s)
    sub, bins = mfdag.tree_pruning(c, w, parent, tol, p)
    oracle = A.tree_pruning(w.tolist(), c.tolist(), tc, tol, p, roots=list(reversed(roots)))
    ratio = check_partition(c, w, parent, tc, sub, bins, tol, p, oracle)
    if ratio <= tol:
        # stopped below tol: the tree ran out of subtrees and the partition
        # is the one for one bin fewer than the failed try, at tol = 0
        nb = int(bins.max()) + 1
        assert mfdag.tree_pruning(c, w, parent, 0.0, nb)[0].tolist() == sub.tolist()
    s.close()


def test_forest_starts_from_every_root():
    systems = [P.anisotropic_laplacian_2d(L, 0.2) for L in (14, 30, 9)]
    s = mfdag.Solver(nb=16, host_only=True)
    s.analyze_batch([(q.ccs_col, q.ccs_row) for q in systems], 8, [q.perm for q in systems])
    c, w, parent, tc, roots = tree_of(s)
    assert len(roots) == 3
    for tol, p in ((0.0, 1), (0.7, 3), (0.9, 6)):
        sub, bins = mfdag.tree_pruning(c, w, parent, tol, p)
        oracle = A.tree_pruning(w.tolist(), c.tolist(), tc, tol, p, roots=list(reversed(roots)))
        check_partition(c, w, parent, tc, sub, bins, tol, p, oracle)
    s.close()


def test_arguments_checked():
    with pytest.raises(mfdag.MfdError):
        mfdag.tree_pruning([1, 1], [1, 0], [0, 1], 0.0, 1)   # parent before child
    with pytest.raises(mfdag.MfdError):
        mfdag.tree_pruning([1, 1], [1, 0], [2, 0], 0.0, 0)   # p < 1
    with pytest.raises(mfdag.MfdError):
        mfdag.tree_pruning([1, 1], [1, 0], [2, 0], 1.0, 1)   # tol >= 1
