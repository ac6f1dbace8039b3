"""Pins the oracle's analysis restatement (oracle/analysis.py) against the
reference's analysis KATs (9x9 fixture) and against invariants the mathematics
fixes: the elimination tree and column counts of a dense symbolic Cholesky,
nested supernode row structures, and a full numeric round trip."""
import numpy as np
import pytest

from oracle import analysis as A


def jag(lengths, vals):
    out, k = [], 0
    for n in lengths:
        out.append(list(vals[k:k + n]))
        k += n
    return out


def test_9x9_function_kats(kats):
    k = kats["analysis_9x9"]
    ccs = jag(k["ccs_num_row"], k["ccs_row"])
    crs = A.ccs_to_crs(ccs)
    assert crs == jag(k["crs_num_col"], k["crs_col"])
    assert A.compute_tree(crs) == k["etree"]
    tc = A.tree_child_from_num_child(k["original_tree_num_child"], k["etree"])
    assert sum(tc, []) == k["original_tree_child_val"]
    assert A.parent_from_tree_child(tc) == k["etree"]
    assert A.subtree_size(tc) == k["original_subtree_size"]
    assert A.postordering(tc) == k["original_postordering"]
    assert A.tree_child_from_parent(k["etree"]) == tc
    assert A.parent_in_postordering_tree(k["postordering_tree_num_child"]) == k["postordering_tree_parent"]
    # column count on the reference's postordered fixture
    pccs = jag(k["postordered_ccs_num_row"], k["postordered_ccs_row"])
    ptc = jag(k["postordering_tree_num_child"], k["postordering_tree_child_val"])
    assert A.column_count(pccs, ptc, k["postordering_tree_parent"]) == k["column_count"]
    assert A.finding_leaves(k["finding_leaves_subtree_size"], pccs) == k["isleaf"]
    lengths = A.node_sets_in_supernode(k["isleaf"], ptc)
    first = A.sets_from_lengths(lengths)
    assert first == k["fundamental_first_node"][: len(first)]
    assert A.supernodal_tree_num_child(first, ptc) == k["supernodal_tree_num_child"]
    sccs = A.supernodal_ccs(first, pccs)
    assert sccs == jag(k["supernodal_ccs_num_row"], k["supernodal_ccs_row"])
    assert A.supernodal_column_count(first, k["column_count"]) == k["supernodal_cc"]
    # repostordering of the original matrix with the reference's postorder perm
    perm, iperm = k["postordering_perm"], k["postordering_iperm"]
    assert A.iperm_of(perm) == iperm
    assert A.repostordering_ccs(ccs, perm, iperm) == pccs


def test_9x9_relax_kats(kats):
    k = kats["analysis_9x9"]
    stc = jag(k["supernodal_tree_num_child"], k["supernodal_tree_child_val"])
    for mz, expected in k["merge_lists"].items():
        merge, _ = A.compute_merge_lists(k["supernodal_cc"], stc, k["supernodal_node_size"], int(mz))
        for node, items in expected.items():
            assert merge.items(int(node)) == items, (mz, node)


def test_9x9_symbolic_local_supernodal_index(kats):
    k = kats["analysis_9x9"]
    stc = jag(k["supernodal_tree_num_child"], k["supernodal_tree_child_val"])
    first = A.sets_from_lengths(k["supernodal_node_size"])
    ccs_a = jag([2, 1, 2, 2, 2, 1, 0], [2, 8, 9, 7, 9, 6, 8, 7, 8, 9])  # symbolic_factorize_test.f90:17-20
    l = A.symbolic_factorize(ccs_a, first, k["supernodal_cc"], stc)
    assert sum(l, []) == k["l_structure_row"]
    lsup = jag(k["supernodal_cc"], k["supernodal_ccs_row"] + [])
    lsup = jag(k["supernodal_ccs_num_row"], k["supernodal_ccs_row"])
    # create_local_index_test.f90 feeds the make_supernodal_ccs rows as L
    lrows = jag(k["supernodal_cc"], k["l_structure_row"])
    assert sum(A.local_index(lrows, first, stc), []) == k["local_index"]
    c = k["supernodal_index_case"]
    sfirst = A.sets_from_lengths(c["node_size"])
    sl = jag(c["l_lengths"], c["l_rows"])
    sa = jag(c["a_num_row"], c["a_rows"])
    assert sum(A.supernodal_index(sfirst, sa, sl), []) == c["expected"]


def random_spd_pattern(n, density, rng):
    """Connected random lower pattern: a path keeps the matrix irreducible."""
    m = rng.random((n, n)) < density
    m = np.tril(m | m.T, -1)
    for i in range(1, n):
        m[i, i - 1] = True
    ccs = []
    for j in range(n):
        rows = [j + 1] + [i + 1 for i in range(j + 1, n) if m[i, j]]
        ccs.append(rows)
    return ccs


def dense_symbolic(ccs, n):
    """Boolean structure of L by dense symbolic elimination (brute force)."""
    S = np.zeros((n, n), bool)
    for j, rows in enumerate(ccs):
        for r in rows:
            S[r - 1, j] = True
    for k in range(n):
        rows = np.nonzero(S[k + 1:, k])[0] + k + 1
        for r in rows:
            S[rows, r] = S[rows, r] | (rows >= r)
    return S


@pytest.mark.parametrize("seed", range(6))
def test_etree_and_column_count_match_dense(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 40))
    ccs = random_spd_pattern(n, 0.12, rng)
    S = dense_symbolic(ccs, n)
    parent = A.compute_tree(A.ccs_to_crs(ccs))
    for j in range(n - 1):
        below = np.nonzero(S[j + 1:, j])[0]
        assert parent[j] == below[0] + j + 2
    tc = A.tree_child_from_parent(parent)
    perm = A.postordering(tc)
    iperm = A.iperm_of(perm)
    pccs = A.repostordering_ccs(ccs, perm, iperm)
    pparent = A.reordering_tree(parent, perm, iperm)
    for i in range(n - 1):
        assert pparent[i] > i + 1
    cc = A.column_count(pccs, A.tree_child_from_parent(pparent), pparent)
    Sp = dense_symbolic(pccs, n)
    assert cc == [int(Sp[:, j].sum()) for j in range(n)]


@pytest.mark.parametrize("seed,max_zero", [(0, 0), (1, 4), (2, 16), (3, 64)])
def test_supernodal_structure_is_exact(seed, max_zero):
    """Row lists of every relaxed supernode equal the union of its columns'
    L-structure below the supernode (dense symbolic Cholesky of P A P^T)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 60))
    ccs = random_spd_pattern(n, 0.08, rng)
    res = A.hppllt_analyze(ccs, max_zero)
    perm = res["perm"]
    iperm = A.iperm_of(perm)
    pccs = [sorted(iperm[r - 1] for r in ccs[perm[j] - 1]) for j in range(n)]
    # symmetric permutation of a lower pattern: rebuild it as lower
    cols = [set() for _ in range(n)]
    for j0, rows in enumerate(ccs):
        for r0 in rows:
            a, b = iperm[j0] , iperm[r0 - 1]
            cols[min(a, b) - 1].add(max(a, b))
    S = dense_symbolic([sorted(c) for c in cols], n)
    first = res["first"]
    for s in range(len(first) - 1):
        lo, hi = first[s] - 1, first[s + 1] - 1
        rows = set()
        for j in range(lo, hi):
            rows |= set((np.nonzero(S[hi:, j])[0] + hi + 1).tolist())
        assert sorted(rows) == res["l_structure"][s]
        if s < len(first) - 2:
            p = res["parent"][s]
            assert p > s + 1
    # every node but the root has a parent; postorder
    assert res["parent"][-1] == 0
    assert all(p == 0 or p > i + 1 for i, p in enumerate(res["parent"]))



def test_block_stencil_generator_is_spd_and_ordered():
    """The 3-unknowns-per-node block stencil: symmetric positive definite,
    lower CCS with the diagonal first, and an ordering that keeps a node's
    unknowns together."""
    from mfdag import problems as P
    p = P.block_stencil_3dof(4, seed=1)
    A = p.dense()
    assert np.allclose(A, A.T)
    assert np.linalg.eigvalsh(A).min() > 0
    assert all(p.ccs_row[p.ccs_col[j] - 1] == j + 1 for j in range(p.n))
    perm = p.perm - 1
    assert sorted(perm) == list(range(p.n))
    assert np.all(perm.reshape(-1, 3) // 3 == (perm.reshape(-1, 3) // 3)[:, :1])


def _nd_lower_ccs(p):
    """Lower CCS (diagonal first) of a generator's matrix in its nested
    dissection order."""
    n = p.n
    perm = p.perm - 1
    iperm = np.empty(n, int)
    iperm[perm] = np.arange(n)
    cols = [set() for _ in range(n)]
    for j in range(n):
        for k in range(p.ccs_col[j] - 1, p.ccs_col[j + 1] - 1):
            a, b = iperm[j], iperm[p.ccs_row[k] - 1]
            cols[min(a, b)].add(max(a, b) + 1)
    return [sorted(c | {j + 1}) for j, c in enumerate(cols)]


@pytest.mark.parametrize("gen,size,max_zero", [("laplacian_2d", 16, 1024), ("laplacian_2d", 20, 128),
                                               ("weighted_laplacian_3d", 8, 512)])
def test_son_list_join_order_is_unobservable(gen, size, max_zero):
    """compute_merge_lists joins the merged children's son lists in stack
    order (relax_supernode_m.f90:52-55); the reference's own test leaves the
    son lists unasserted (relax_supernode_test.f90:2).  The order cannot
    reach any output, since the lists are afterwards read only for their
    lengths: joining them in the opposite order gives the same analysis.
    The cases are ones where the two orders give differently ordered lists."""
    from mfdag import problems as P
    ccs = _nd_lower_ccs(getattr(P, gen)(size))
    a = A.analyze_phase(ccs, max_zero)
    b = A.analyze_phase(ccs, max_zero, son_merge_fifo=True)
    assert a.keys() == b.keys()
    for key in a:
        assert a[key] == b[key], key
    fa = a["fundamental_lengths"]
    m1, s1 = A.compute_merge_lists(a["fundamental_cc"], a["fundamental_tree_child"], fa, max_zero)
    m2, s2 = A.compute_merge_lists(a["fundamental_cc"], a["fundamental_tree_child"], fa, max_zero, True)
    differ = 0
    for i in range(1, len(fa) + 1):
        assert sorted(s1.items(i)) == sorted(s2.items(i))
        assert m1.items(i) == m2.items(i)
        differ += s1.items(i) != s2.items(i)
    assert differ > 0
