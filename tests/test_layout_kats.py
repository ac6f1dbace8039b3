"""Pins the oracle's layout and structure restatement against the
reference's own known answers (tests/golden/layout_kats.json, transcribed
from the reference's tests by tests/golden/transcribe_layout_kats.py):

  * where each tile of the supernode / work / border factor arrays and each
    block of the supernode / work / border right-hand-side arrays lives, and
    every node's share of each array (factors/*_extractor_test.f90,
    right_hand/*_array_extractor_test.f90);
  * the node_data quantities (node_data_test.f90) and get_block_size
    (block_size_calculator_test.f90);
  * the block_local_index groups and their tile-local rows
    (block_local_index_creator_test.f90);
  * build_map / create_node_sets / create_perm / build_cc / count_num_child
    (relaxed_supernode_test.f90), the permutations (perm_test.f90) and the
    reorderings (reordering_test.f90).

The order inside the relaxed son lists is not pinned and needs no pin:
count_num_child reads only their lengths, and no other output reads them."""
import json
import os

import pytest

from oracle import analysis as A
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lk():
    with open(os.path.join(HERE, "golden", "layout_kats.json")) as f:
        return json.load(f)


def oracle_nodes(c, w, nb):
    m = len(c)
    parent = [k + 2 for k in range(m - 1)] + [0]
    return O.OracleFactor(c, w, parent, nb, [1] * max(sum(w), 1))


MATRIX = {"supernode_matrix": (1, 10, 13), "work_matrix": (2, 11, 14), "border_matrix": (3, 12, 15)}
ARRAY = {"supernode_array": (16, 19), "work_array": (17, 20), "border_array": (18, 21)}


@pytest.mark.parametrize("kind", sorted(MATRIX))
def test_factor_array_extractors(lk, kind):
    which, size_q, pos_q = MATRIX[kind]
    cases = lk["extractors"][kind]["cases"]
    assert cases
    for case in cases:
        f = oracle_nodes(case["c"], case["w"], case["nb"])
        for node, size in case["sizes"].items():
            assert f.layout(size_q, int(node)) == size, (kind, case["nb"], node)
        for node, i, j, first, last in case["ptrs"]:
            start = f.layout(pos_q, node, i, j)
            _, rows, cols = f.tile(which, node, i, j)
            assert (start + 1, start + rows * cols) == (first, last), (kind, node, i, j)


@pytest.mark.parametrize("kind", sorted(ARRAY))
def test_rhs_array_extractors(lk, kind):
    size_q, pos_q = ARRAY[kind]
    for case in lk["extractors"][kind]["cases"]:
        f = oracle_nodes(case["c"], case["w"], case["nb"])
        for node, size in case["sizes"].items():
            assert f.layout(size_q, int(node)) == size, (kind, case["nb"], node)
        for node, idx, first, last in case["ptrs"]:
            start = f.layout(pos_q, node, idx)
            if kind == "supernode_array":
                length = f.layout(6, node, idx)
            elif kind == "work_array":
                length = f.layout(7, node, idx)
            else:
                length = f.layout(size_q, node)
            assert (start + 1, start + length) == (first, last), (kind, node, idx)


NODE_ARRAYS = {"divisible": 0, "border_supernode_size": 1, "border_work_size": 2, "num_matrix_block": 3,
               "work_start_index": 4}
NODE_CALLS = {"get_num_matrix_block": 3, "get_work_start_index": 4, "get_matrix_block_size": 5,
              "get_supernode_block_size": 6, "get_work_block_size": 7, "get_matrix_num": 8, "get_work_num": 9}


def test_node_data(lk):
    cases = lk["node_data"]["cases"]
    assert len(cases) == 2
    for case in cases:
        f = oracle_nodes(case["c"], case["w"], case["nb"])
        for name, vals in case["arrays"].items():
            got = [f.layout(NODE_ARRAYS[name], s) for s in range(1, len(vals) + 1)]
            assert got == vals, name
        for call in case["calls"]:
            name, args, want = call[0], call[1:-1], call[-1]
            q = NODE_CALLS[name]
            if q in (3, 4):
                got = f.layout(q, args[0])
            elif q == 8:
                got = f.layout(q, 1, args[0])
            else:
                got = f.layout(q, args[1], args[0])
            assert got == want, call
    assert sum(len(c["calls"]) for c in cases) == 93


def test_block_size_calculator(lk):
    for idx, nb, n, first, want in lk["block_size_calculator"]["cases"]:
        assert O.lib().orc_block_size(idx, nb, n, first or 0) == want, (idx, nb, n, first)


def test_block_local_index_creator(lk):
    for case in lk["block_local_index_creator"]["cases"]:
        f = O.OracleFactor(case["c"], case["w"], [k + 2 for k in range(len(case["c"]) - 1)] + [0], case["nb"],
                           case["local_index"])
        nodes = len(case["num_blocks"])
        groups = [f.groups(s) for s in range(1, nodes + 1)]
        assert [len(g) for g in groups] == case["num_blocks"]
        assert [len(g[0]) for gs in groups for g in gs] == case["num_indices"]
        rebuilt = [v for gs in groups for g in gs for v in g[0]]
        assert rebuilt == case["rebuild_val"][:len(rebuilt)]
        assert len(rebuilt) == sum(case["w"])


def _lists(n, merges):
    ml = A.LinkedLists(n)
    for i in range(1, n + 1):
        ml.add(i, i)
    for frm, to in merges:
        ml.merge(frm, to)
    return ml


def test_relaxed_supernode(lk):
    r = lk["relaxed_supernode"]
    for case in r["cases"]:
        merge = _lists(7, case["merges"])
        fmap = A.build_map(merge)
        assert fmap == case["map"]
        lengths = A.create_node_sets(r["node_size"], fmap, merge)
        assert A.sets_from_lengths(lengths) == case["first_node"]
        assert A.create_perm(r["node_size"], fmap, merge) == case["perm"]
        assert A.build_cc(r["cc_fundamental"], fmap) == case["cc"]
        sons = A.LinkedLists(7)
        for node, idx in case["sons"]:
            sons.add(node, idx)
        assert A.count_num_child(sons, fmap) == case["num_child"]


def test_perm_and_reordering(lk, kats):
    p = lk["perm"]
    perm, iperm = p["perm"], p["iperm"]
    assert A.permutate(perm, list(range(1, 10))) == p["permutate_of_1_to_9"]
    assert A.inverse_permutate(perm, list(range(1, 10))) == p["inverse_permutate_of_1_to_9"]
    assert A.iperm_of(perm) == iperm
    k = kats["analysis_9x9"]
    ccs, off = [], 0
    for ln in k["ccs_num_row"]:
        ccs.append(k["ccs_row"][off:off + ln])
        off += ln
    r = lk["reordering"]
    out, cperm = A.repostordering_ccs(ccs, perm, iperm, with_perm=True)
    assert cperm == r["ccs_perm_postorder"]
    vals = A.reordering_ccs_val([len(c) for c in ccs], [len(c) for c in out], r["origin_val"], perm)
    assert vals == r["ordered_val"]
    assert A.permutate(cperm, r["origin_val"]) == r["ordered_val"]
    sm = r["small"]
    sccs, off = [], 0
    for ln in sm["num_row"]:
        sccs.append(sm["row"][off:off + ln])
        off += ln
    out, cp = A.reordering_ccs(sccs, sm["perm"], sm["iperm"])
    assert [len(c) for c in out] == sm["out_num_row"]
    assert sum(out, []) == sm["out_row"]
    assert cp == sm["ccs_perm"]
