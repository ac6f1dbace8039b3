"""Pure-Python restatement of the reference's host symbolic analysis.

TEST INFRASTRUCTURE ONLY: this is the oracle for the product's C++ analysis
(multifrontal-dag_amd/csrc/mfd_analysis.cpp).  Only tests/, the smoke check
and bench.py's cpu_baseline leg may import it, and only as the checker.

Every function follows one reference routine statement by statement (cited
per function, paths relative to the reference root).  All indices are 1-based
as in the reference; a "jagged" array is a list of Python lists.  Plain loops
are used on purpose: the checker must be easy to audit, and it is only run on
matrices of up to ~10^4 columns.
"""
from __future__ import annotations


# --------------------------------------------------------------- containers
class DisjointSet:
    """src/main/data_structures/disjoint_set_m.f90:16-69 (real/virtual roots)."""

    def __init__(self, n):
        self.parent = list(range(n + 1))
        self.realroot = list(range(n + 1))
        self.virtualroot = list(range(n + 1))
        self.rank = [0] * (n + 1)

    def find2(self, idx):
        # iterative form of the recursive path-compressing find2 (:32-38)
        root = idx
        while self.parent[root] != root:
            root = self.parent[root]
        while self.parent[idx] != root:
            nxt = self.parent[idx]
            self.parent[idx] = root
            idx = nxt
        return root

    def find(self, idx):
        return self.realroot[self.find2(idx)]

    def link(self, node, parent):
        nv = self.virtualroot[node]
        pv = self.virtualroot[parent]
        if self.rank[parent] > self.rank[node]:
            self.parent[nv] = pv
        elif self.rank[parent] < self.rank[node]:
            self.parent[pv] = nv
            self.realroot[nv] = self.realroot[pv]
            self.virtualroot[parent] = nv
        else:
            self.parent[nv] = pv
            self.rank[parent] += 1


class LinkedLists:
    """src/main/data_structures/doubly_linked_lists_m.f90:21-133."""

    def __init__(self, n):
        self.head = [0] * (n + 1)
        self.tail = [0] * (n + 1)
        self.next = [0] * (n + 1)
        self.prev = [0] * (n + 1)

    def add(self, node, idx):
        if self.head[idx] == 0:
            self.head[idx] = node
            self.tail[idx] = node
        else:
            self.prev[node] = self.tail[idx]
            self.next[self.tail[idx]] = node
            self.tail[idx] = node

    def merge(self, from_list, to_list):
        if self.head[from_list] == 0:
            return
        if self.head[to_list] == 0:
            self.head[to_list] = self.head[from_list]
            self.tail[to_list] = self.tail[from_list]
        else:
            self.next[self.tail[to_list]] = self.head[from_list]
            self.tail[to_list] = self.tail[from_list]
        self.head[from_list] = 0
        self.tail[from_list] = 0

    def remove(self, node, idx):
        nxt = self.next[node]
        prv = self.prev[node]
        if node == self.head[idx]:
            self.head[idx] = nxt
        else:
            self.next[prv] = nxt
        if node == self.tail[idx]:
            self.tail[idx] = prv
        else:
            self.prev[nxt] = prv
        self.next[node] = 0
        self.prev[node] = 0

    def items(self, idx):
        out = []
        node = self.head[idx]
        while node != 0:
            out.append(node)
            node = self.next[node]
        return out

    def num_elements(self):
        return len(self.head) - 1


def sets_from_lengths(lengths):
    """contiguous_sets_m.f90:18-30: returns 1-based first indices (len+1)."""
    ptr = [1]
    for ln in lengths:
        ptr.append(ptr[-1] + ln)
    return ptr


# --------------------------------------------------------------- algorithms
def ccs_to_crs(ccs):
    """src/main/storage_converter/ccs_to_crs_m.f90:9-43."""
    n = len(ccs)
    crs = [[] for _ in range(n)]
    for j in range(1, n + 1):
        for r in ccs[j - 1]:
            crs[r - 1].append(j)
    return crs


def compute_tree(crs):
    """src/main/algorithms/compute_tree_m.f90:10-35."""
    n = len(crs)
    parent = [0] * (n + 1)
    ds = DisjointSet(n)
    for i in range(1, n + 1):
        parent[i] = 0
        cols = crs[i - 1]
        for k in range(len(cols) - 1):
            t = ds.find(cols[k])
            if parent[t] == 0 and t != i:
                parent[t] = i
                ds.link(t, i)
    return parent[1:]


def tree_child_from_num_child(num_child, parent):
    """tree_m.f90:15-37 (create_tree_child_by_num_child_and_parent)."""
    n = len(num_child)
    child = [[] for _ in range(n)]
    for i in range(1, n + 1):
        if parent[i - 1] == 0:
            continue
        child[parent[i - 1] - 1].append(i)
    for i in range(n):
        assert len(child[i]) == num_child[i]
    return child


def tree_child_from_parent(parent):
    """tree_m.f90:39-55 (create_tree_child_by_parent; the root is node n)."""
    n = len(parent)
    num_child = [0] * n
    for i in range(1, n):
        num_child[parent[i - 1] - 1] += 1
    return tree_child_from_num_child(num_child, parent)


def parent_in_postordering_tree(num_child):
    """tree_m.f90:57-76."""
    n = len(num_child)
    stack = []
    parent = [0] * n
    for i in range(1, n + 1):
        for _ in range(num_child[i - 1]):
            node = stack.pop()
            parent[node - 1] = i
        stack.append(i)
    return parent


def postordering(tree_child):
    """tree_m.f90:78-99 (tree_traverse_postordering)."""
    n = len(tree_child)
    ptr = n
    perm = [0] * n
    stack = [n]
    while stack:
        node = stack.pop()
        perm[ptr - 1] = node
        ptr -= 1
        for c in tree_child[node - 1]:
            stack.append(c)
    return perm


def subtree_size(tree_child):
    """tree_m.f90:101-116."""
    n = len(tree_child)
    size = [0] * n
    for i in range(1, n + 1):
        size[i - 1] = 1 + sum(size[c - 1] for c in tree_child[i - 1])
    return size


def parent_from_tree_child(tree_child):
    """tree_m.f90:118-132 (create_parent)."""
    n = len(tree_child)
    parent = [0] * n
    for node in range(1, n + 1):
        for c in tree_child[node - 1]:
            parent[c - 1] = node
    return parent


def iperm_of(perm):
    """perm_m.f90:11-23."""
    ip = [0] * len(perm)
    for i, p in enumerate(perm, start=1):
        ip[p - 1] = i
    return ip


def permutate(perm, data):
    """perm_m.f90 permutate_DP / permutate_int: out(i) = in(perm(i))."""
    return [data[p - 1] for p in perm]


def inverse_permutate(perm, data):
    """perm_m.f90 inverse_permutate: out(perm(i)) = in(i)."""
    out = [None] * len(data)
    for i, p in enumerate(perm):
        out[p - 1] = data[i]
    return out


def perm_product(first, second):
    """perm_m.f90:25-36: out(i) = first(second(i))."""
    return [first[s - 1] for s in second]


def repostordering_ccs(ccs, perm, iperm, with_perm=False):
    """reordering_m.f90:31-96 (repostordering_ccs1 / repostordering_ccs2):
    column perm(c) becomes column c, rows are renumbered but not re-sorted."""
    n = len(ccs)
    out = [None] * n
    offsets = []
    off = 0
    for j in range(n):
        offsets.append(off)
        off += len(ccs[j])
    ccs_perm = [None] * n
    for col_o in range(1, n + 1):
        col_r = iperm[col_o - 1]
        out[col_r - 1] = [iperm[r - 1] for r in ccs[col_o - 1]]
        ccs_perm[col_r - 1] = [offsets[col_o - 1] + i for i in range(1, len(ccs[col_o - 1]) + 1)]
    if with_perm:
        return out, [x for col in ccs_perm for x in col]
    return out


def reordering_ccs(ccs, perm, iperm):
    """reordering_m.f90:98-148: symmetric permutation of a lower CCS; rows of
    each new column are sorted (with their source positions)."""
    n = len(ccs)
    cols = [[] for _ in range(n)]
    off = 0
    for j in range(1, n + 1):
        rows = ccs[j - 1]
        for i, r in enumerate(rows, start=1):
            c = min(iperm[j - 1], iperm[r - 1])
            rr = max(iperm[j - 1], iperm[r - 1])
            cols[c - 1].append((rr, off + i))
        off += len(rows)
    out, cperm = [], []
    for col in cols:
        # sort_with_perm is a stable bubble sort on distinct keys
        col.sort(key=lambda t: t[0])
        out.append([t[0] for t in col])
        cperm.extend(t[1] for t in col)
    return out, cperm


def reordering_ccs_val(origin_lengths, reorder_lengths, origin_val, perm):
    """reordering_m.f90:150-170: column perm(c) of the values moves, whole,
    to column c."""
    of = sets_from_lengths(origin_lengths)
    out = []
    for c in range(1, len(reorder_lengths) + 1):
        o = perm[c - 1]
        out.extend(origin_val[of[o - 1] - 1:of[o] - 1])
    return out


def reordering_tree(parent, perm, iperm):
    """reordering_m.f90:17-29."""
    n = len(parent)
    out = [0] * n
    for i in range(1, n):
        out[i - 1] = iperm[parent[perm[i - 1] - 1] - 1]
    out[n - 1] = 0
    return out


def column_count(ccs, tree_child, parent):
    """src/main/algorithms/column_count_m.f90:9-72."""
    n = len(parent)
    ds = DisjointSet(n)
    lca = [[0] * len(col) for col in ccs]
    prev_p = [0] * (n + 1)
    for j in range(1, n):
        rows = ccs[j - 1]
        for i in range(1, len(rows)):
            r = rows[i]
            pp = prev_p[r]
            lca[j - 1][i] = ds.find(pp) if pp != 0 else 0
            prev_p[r] = j
        ds.link(j, parent[j - 1])
    wt = [0] * (n + 1)
    for i in range(1, n + 1):
        wt[i] = 1 if len(tree_child[i - 1]) == 0 else 0
    for j in range(1, n + 1):
        if j != n:
            wt[parent[j - 1]] -= 1
        rows = ccs[j - 1]
        for i in range(1, len(rows)):
            wt[j] += 1
            if lca[j - 1][i] != 0:
                wt[lca[j - 1][i]] -= 1
    cc = wt[1:]
    for i in range(1, n):
        cc[parent[i - 1] - 1] += cc[i - 1]
    return cc


def finding_leaves(sub_size, ccs):
    """src/main/algorithms/finding_leaves_m.f90:6-29."""
    n = len(ccs)
    isleaf = [0] * n
    prev = [0] * (n + 1)
    for j in range(1, n + 1):
        rows = ccs[j - 1]
        for i in range(1, len(rows)):
            r = rows[i]
            if prev[r] < j - sub_size[j - 1] + 1:
                isleaf[j - 1] = 1
            prev[r] = j
    return isleaf


def node_sets_in_supernode(isleaf, tree_child):
    """fundamental_supernode_m.f90:11-38: returns supernode lengths."""
    lengths = []
    for i in range(1, len(tree_child) + 1):
        if isleaf[i - 1] > 0 or len(tree_child[i - 1]) > 1:
            lengths.append(0)
        lengths[-1] += 1
    return lengths


def supernodal_tree_num_child(first, tree_child):
    """fundamental_supernode_m.f90:55-67."""
    return [len(tree_child[f - 1]) for f in first[:-1]]


def supernodal_column_count(first, cc_node):
    """fundamental_supernode_m.f90:40-53."""
    return [cc_node[first[i + 1] - 2] - 1 for i in range(len(first) - 1)]


def supernodal_ccs(first, ccs):
    """fundamental_supernode_m.f90:70-117 (create_supernodal_ccs)."""
    ns = len(first) - 1
    out = []
    for k in range(1, ns + 1):
        last = first[k] - 1
        seen = set()
        rows_k = []
        for j in range(first[k - 1], first[k]):
            for r in ccs[j - 1]:
                if r <= last or r in seen:
                    continue
                seen.add(r)
                rows_k.append(r)
        out.append(rows_k)
    return out


def compute_merge_lists(cc, tree_child, lengths, max_zero, son_merge_fifo=False):
    """src/main/algorithms/relax_supernode_m.f90:13-90.

    The son lists of the merged children are joined to node k's list in the
    order the stack pops them (:52-55).  That order cannot change any output:
    choose_merge_node (:66-85) only walks the list of the node being
    processed, and only before the joins of that node; afterwards the lists
    are read only for their lengths (count_num_child,
    relaxed_supernode_m.f90:87-99).  son_merge_fifo=True joins them in the
    other order; the tests use it to show that the analysis is the same."""
    n = len(cc)
    sons = LinkedLists(n)
    merge = LinkedLists(n)
    for i in range(1, n + 1):
        merge.add(i, i)
    num_cols = [0] + list(lengths)
    num_zeros = [0] * (n + 1)
    stack = []

    def additional(p, c):
        return (cc[p - 1] + num_cols[p] - cc[c - 1]) * num_cols[c]

    for k in range(1, n + 1):
        if len(tree_child[k - 1]) == 0:
            continue
        for c in tree_child[k - 1]:
            sons.add(c, k)
        while True:
            # choose_merge_node (:66-85): the last son with the fewest zeros
            min_zero = max_zero
            merge_node = 0
            for c in sons.items(k):
                z = additional(k, c) + num_zeros[k] + num_zeros[c]
                if z <= min_zero and z <= max_zero:
                    merge_node = c
                    min_zero = z
            if merge_node != 0:
                merge.merge(merge_node, k)
                sons.remove(merge_node, k)
                stack.append(merge_node)
                num_zeros[k] = min_zero
                num_cols[k] += num_cols[merge_node]
            else:
                while stack:
                    sons.merge(stack.pop(0) if son_merge_fifo else stack.pop(), k)
                break
    return merge, sons


def build_map(merge):
    """relaxed_supernode_m.f90:15-31: the lists still holding nodes, in index
    order (each is one relaxed supernode)."""
    return [i for i in range(1, merge.num_elements() + 1) if merge.head[i] != 0]


def create_node_sets(fund_lengths, fmap, merge):
    """relaxed_supernode_m.f90:64-86: columns of each relaxed supernode."""
    return [sum(fund_lengths[x - 1] for x in merge.items(m)) for m in fmap]


def create_perm(fund_lengths, fmap, merge):
    """relaxed_supernode_m.f90:33-62: each merge list pushed onto a stack in
    list order and popped (last pushed first), each popped fundamental
    supernode contributing its columns in order."""
    fund_first = sets_from_lengths(fund_lengths)
    perm = []
    for m in fmap:
        stack = []
        for node in merge.items(m):
            stack.append(node)
        while stack:
            node = stack.pop()
            perm.extend(range(fund_first[node - 1], fund_first[node]))
    return perm


def count_num_child(sons, fmap):
    """relaxed_supernode_m.f90:88-99: only the son lists' lengths are used."""
    return [len(sons.items(m)) for m in fmap]


def build_cc(fund_cc, fmap):
    """relaxed_supernode_m.f90:101-112."""
    return [fund_cc[m - 1] for m in fmap]


def relax_supernode(fund_lengths, fund_cc, fund_tree_child, fund_ccs, max_zero, son_merge_fifo=False):
    """analyse_phase_m.f90:91-108 with relaxed_supernode_m.f90:15-112."""
    merge, sons = compute_merge_lists(fund_cc, fund_tree_child, fund_lengths, max_zero, son_merge_fifo)
    fmap = build_map(merge)
    rel_lengths = create_node_sets(fund_lengths, fmap, merge)
    perm = create_perm(fund_lengths, fmap, merge)
    num_child = count_num_child(sons, fmap)
    parent = parent_in_postordering_tree(num_child)
    tree_child = tree_child_from_num_child(num_child, parent)
    cc = build_cc(fund_cc, fmap)
    iperm = iperm_of(perm)
    ccs = repostordering_ccs(fund_ccs, perm, iperm)
    return dict(lengths=rel_lengths, perm=perm, tree_child=tree_child, cc=cc, ccs=ccs,
                merge_lists={m: merge.items(m) for m in fmap})


def symbolic_factorize(ccs_a, first, cc, tree_child):
    """src/main/algorithms/symbolic_factorize_m.f90:10-62."""
    n = len(ccs_a)
    order = first[-1] - 1
    full = [0] * (order + 1)
    rows_l = [None] * n
    for i in range(1, n + 1):
        rows = []
        for r in ccs_a[i - 1]:
            rows.append(r)
            full[r] = i
        last = first[i] - 1
        for c in tree_child[i - 1]:
            for r in rows_l[c - 1]:
                if r <= last or full[r] == i:
                    continue
                rows.append(r)
                full[r] = i
        assert len(rows) == cc[i - 1], "column count disagrees with symbolic structure"
        rows_l[i - 1] = rows
    return [sorted(r) for r in rows_l]


def analyze_phase(ccs, max_zero, son_merge_fifo=False):
    """src/main/analyse_phase_m.f90:22-89."""
    n = len(ccs)
    crs = ccs_to_crs(ccs)
    parent = compute_tree(crs)
    tree_child = tree_child_from_parent(parent)
    pperm = postordering(tree_child)
    piperm = iperm_of(pperm)
    pccs = repostordering_ccs(ccs, pperm, piperm)
    pparent = reordering_tree(parent, pperm, piperm)
    ptree_child = tree_child_from_parent(pparent)
    cc_node = column_count(pccs, ptree_child, pparent)
    # finding_supernode (:73-89)
    ssize = subtree_size(ptree_child)
    isleaf = finding_leaves(ssize, pccs)
    flen = node_sets_in_supernode(isleaf, ptree_child)
    ffirst = sets_from_lengths(flen)
    fnum_child = supernodal_tree_num_child(ffirst, ptree_child)
    fparent = parent_in_postordering_tree(fnum_child)
    ftree_child = tree_child_from_num_child(fnum_child, fparent)
    fcc = supernodal_column_count(ffirst, cc_node)
    rel = relax_supernode(flen, fcc, ftree_child, pccs, max_zero, son_merge_fifo)
    rfirst = sets_from_lengths(rel["lengths"])
    sccs = supernodal_ccs(rfirst, rel["ccs"])
    l_structure = symbolic_factorize(sccs, rfirst, rel["cc"], rel["tree_child"])
    perm = perm_product(pperm, rel["perm"])
    return dict(
        etree=parent, postorder=pperm, column_count=cc_node, isleaf=isleaf,
        fundamental_lengths=flen, fundamental_cc=fcc, fundamental_tree_child=ftree_child,
        merge_lists=rel["merge_lists"], node_sizes=rel["lengths"], l_structure=l_structure,
        perm=perm, parent=parent_from_tree_child(rel["tree_child"]), tree_child=rel["tree_child"],
        cc=rel["cc"])


def supernodal_index(first, a_structure, l_structure):
    """src/main/algorithms/create_supernodal_index_m.f90:10-40."""
    order = first[-1] - 1
    mp = [0] * (order + 1)
    out = []
    for node in range(1, len(first)):
        off = first[node - 1] - 1
        nc = first[node] - first[node - 1]
        for j in range(1, nc + 1):
            mp[off + j] = j
        for i, r in enumerate(l_structure[node - 1], start=1):
            mp[r] = nc + i
        for j in range(first[node - 1], first[node]):
            out.append([mp[r] for r in a_structure[j - 1]])
    return out


def local_index(l_structure, first, tree_child):
    """src/main/algorithms/create_local_index_m.f90:10-45 (before the in-place
    mod2 of block_local_index_info_m.f90:54).  Roots keep zeros."""
    order = first[-1] - 1
    mp = [0] * (order + 1)
    out = [[0] * len(r) for r in l_structure]
    for node in range(1, len(l_structure) + 1):
        if len(tree_child[node - 1]) == 0:
            continue
        nc = first[node] - first[node - 1]
        off = first[node - 1] - 1
        for i in range(1, nc + 1):
            mp[off + i] = i
        for i, r in enumerate(l_structure[node - 1], start=1):
            mp[r] = nc + i
        for c in tree_child[node - 1]:
            out[c - 1] = [mp[r] for r in l_structure[c - 1]]
    return out


def hppllt_analyze(ccs, max_zero, perm_in=None):
    """hppllt.f90:24-106 (hppllt_analyze + create_data_structure, integer part).
    `perm_in` (new -> old, 1-based) takes the place of the METIS ordering; None
    means use_metis = 0."""
    if perm_in is not None:
        iperm_in = iperm_of(perm_in)
        reordered, reorder_ccs_perm = reordering_ccs(ccs, perm_in, iperm_in)
    else:
        reordered, reorder_ccs_perm = ccs, None
    res = analyze_phase(reordered, max_zero)
    aperm = res["perm"]
    aiperm = iperm_of(aperm)
    a_structure, a_ccs_perm = repostordering_ccs(reordered, aperm, aiperm, with_perm=True)
    if perm_in is not None:
        perm = perm_product(perm_in, aperm)
        ccs_perm = perm_product(reorder_ccs_perm, a_ccs_perm)
    else:
        perm = aperm
        ccs_perm = a_ccs_perm
    first = sets_from_lengths(res["node_sizes"])
    res.update(
        perm=perm, ccs_perm=ccs_perm, a_structure=a_structure, first=first,
        work_sizes=[len(r) for r in res["l_structure"]],
        supernodal_index=supernodal_index(first, a_structure, res["l_structure"]),
        local_index=local_index(res["l_structure"], first, res["tree_child"]))
    return res


def cost(node_sizes, work_sizes):
    """cost_calculator_m.f90:9-20."""
    return sum(sum(m * m for m in range(w + 1, w + c + 1)) for c, w in zip(node_sizes, work_sizes))


def nonzero(node_sizes, work_sizes):
    """cost_calculator_m.f90:22-33."""
    return sum(sum(range(w + 1, w + c + 1)) for c, w in zip(node_sizes, work_sizes))


# ------------------------------------------------------------ tree pruning
def count_op_count(cc, node_sizes):
    """tree_pruning_m.f90:76-88: sum of m^2 for m = cc+1 .. cc+size."""
    out = []
    for l, size in zip(cc, node_sizes):
        k = size + l
        out.append(k * (k + 1) * (2 * k + 1) // 6 - l * (l + 1) * (2 * l + 1) // 6)
    return out


def generate_tree_weights(fcnt, tree_child):
    """tree_pruning_m.f90:90-110 (children precede parents)."""
    nodewt = []
    for i, f in enumerate(fcnt):
        nodewt.append(f + sum(nodewt[ch - 1] for ch in tree_child[i]))
    return nodewt


def partition_tree(lst, nodewt, tree_child, ip):
    """tree_pruning_m.f90:112-134 with doubly_linked_list_m's add (at the
    head) and find_node_with_max_op (:136-154, the first of the largest).
    Returns False when the list runs out."""
    while len(lst) < ip:
        node, best = 0, 0
        for v in lst:
            if nodewt[v - 1] > best:
                node, best = v, nodewt[v - 1]
        if node == 0:
            return False
        for ch in tree_child[node - 1]:
            lst.insert(0, ch)
        lst.remove(node)
    return True


def tree_pruning(cc, node_sizes, tree_child, tol, p, roots=None):
    """tree_pruning_m.f90:13-74: cc = work sizes, tree_child 1-based jagged.
    roots: the list to start from (the reference starts from node n)."""
    fcnt = count_op_count(cc, node_sizes)
    nodewt = generate_tree_weights(fcnt, tree_child)
    n = len(cc)
    lst = list(roots) if roots is not None else [n]
    ratio = 0.0
    ip = p - 1
    while ratio <= tol:
        ip += 1
        if not partition_tree(lst, nodewt, tree_child, ip):
            return tree_pruning(cc, node_sizes, tree_child, 0.0, ip - 1, roots)
        # op_heap (largest first) into bin_heap (least loaded first)
        bins = [0] * ip
        for wgt in sorted((nodewt[v - 1] for v in lst), reverse=True):
            k = bins.index(min(bins))
            bins[k] += wgt
        ratio = min(bins) / max(bins)
    return lst
