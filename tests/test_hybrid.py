"""CPU+GPU hybrid: subtrees of the partition (mfd_tree_pruning, the
reference's tree_pruning, tree_pruning_m.f90:13-73) factored on the host
(mfd_host.cpp, the front loop of seq_factorize_m.f90:13-98) while the GPU
factors the rest (mfd_set_host_subtrees).

CPU: the host part alone against the oracle's fronts (1e-12 relative per
front) on analysed problems, a pivot that is not positive inside a host
subtree reported at the oracle's column, and the argument checks (roots out
of range, repeated or nested).
GPU: the hybrid factorization's fronts (host and device) against the
oracle's (1e-12), its solution against the GPU-only one (1e-12) with
eta <= 1e-14 for 1 and 3 right-hand sides, the device-values entry point, a
second factorization on the same handle, and the forms the hybrid refuses."""
import sys

import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from parity import front_errors, oracle_for


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host_roots(solver, p=4, tol=0.5, bin_=0):
    """0-based roots of the subtrees the partition deals to bin bin_."""
    st = solver.structure()
    sub, bins = mfdag.tree_pruning(st["supernode_size"], st["work_size"], st["parent"], tol, p)
    return np.sort(sub[bins == bin_] - 1)


def subtree_nodes(solver, roots):
    """1-based nodes of the subtrees rooted at the 0-based roots."""
    parent = solver.structure()["parent"]
    m = len(parent)
    host = np.zeros(m, bool)
    host[np.asarray(roots)] = True
    for s in range(m - 1, -1, -1):  # parents follow their children
        if parent[s] > 0 and host[parent[s] - 1]:
            host[s] = True
    return np.flatnonzero(host) + 1


PROBLEMS = {
    "lap2d_40": (lambda: P.laplacian_2d(40), 16, 16),
    "wlap3d_10": (lambda: P.weighted_laplacian_3d(10, seed=1), 64, 32),
    "blk3_6": (lambda: P.block_stencil_3dof(6, seed=7), 64, 32),
}


# ------------------------------------------------------------------- CPU
@pytest.mark.parametrize("name", sorted(PROBLEMS))
def test_host_part_matches_oracle(name):
    make, max_zero, nb = PROBLEMS[name]
    p = make()
    s = mfdag.Solver(nb=nb, host_only=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    roots = host_roots(s)
    assert len(roots) > 0
    s.set_host_subtrees(roots, threads=3)
    assert s.factorize_host_subtrees(p.val) == 0
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    nodes = subtree_nodes(s, roots)
    worst = 0.0
    for node in nodes:
        ref = np.tril(f.front(node))
        got = np.tril(s.front(node))
        worst = max(worst, np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
    ms, nfr, flops = s.host_part()
    log(f"{name}: {len(roots)} subtrees, {nfr} host fronts of {len(st['parent'])}, {flops:.3e} flops, "
        f"{ms:.2f} ms; worst front error {worst:.2e}")
    assert nfr == len(nodes)
    assert worst <= 1e-12
    # a front outside the host subtrees has no factor without a GPU
    others = np.setdiff1d(np.arange(1, len(st["parent"]) + 1), nodes)
    with pytest.raises(mfdag.MfdError):
        s.front(int(others[0]))
    s.close()


def test_host_part_reports_failing_pivot():
    p = P.laplacian_2d(30)
    s = mfdag.Solver(nb=16, host_only=True)
    s.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    roots = host_roots(s)
    st = s.structure()
    # a diagonal entry of the first host front's first pivot column made negative
    node = subtree_nodes(s, roots)[0]
    first = int(np.sum(st["supernode_size"][:node - 1]))  # 0-based postordered column
    a_colptr, snidx, ccs_perm = st["a_colptr"], st["supernodal_index"], st["ccs_perm"]
    val = p.val.copy()
    for e in range(a_colptr[first], a_colptr[first + 1]):
        if snidx[e] == 1:  # the diagonal: front row 1 of the front's first column
            val[ccs_perm[e] - 1] = -4.0
    s.set_host_subtrees(roots, threads=2)
    col = s.factorize_host_subtrees(val)
    _, _, info = oracle_for(s, val)
    log(f"host part failing column {col}, oracle {info}")
    assert col == first + 1
    assert info == col
    s.close()


def test_host_subtree_arguments():
    p = P.weighted_laplacian_3d(8, seed=2)
    s = mfdag.Solver(nb=32, host_only=True)
    s.analyze(p.ccs_col, p.ccs_row, 64, p.perm)
    parent = s.structure()["parent"]
    m = len(parent)
    child = int(np.flatnonzero(parent > 0)[0])
    with pytest.raises(mfdag.MfdError):
        s.set_host_subtrees([m])  # out of range
    with pytest.raises(mfdag.MfdError):
        s.set_host_subtrees([child, child])  # twice
    with pytest.raises(mfdag.MfdError):
        s.set_host_subtrees([child, int(parent[child]) - 1])  # nested
    with pytest.raises(mfdag.MfdError):
        s.set_host_subtrees([child], threads=0)
    with pytest.raises(mfdag.MfdError):
        s.factorize_host_subtrees(p.val)  # no host subtrees set
    s.set_host_subtrees([child])
    assert s.factorize_host_subtrees(p.val) == 0
    s.set_host_subtrees([])
    with pytest.raises(mfdag.MfdError):
        s.factorize_host_subtrees(p.val)
    s.close()


# ------------------------------------------------------------------- GPU
GPU_CASES = [
    ("wlap3d_20", lambda: P.weighted_laplacian_3d(20, seed=3), 64, 32, 4),
    ("blk3_12", lambda: P.block_stencil_3dof(12, seed=7), 64, 32, 8),
    ("lap2d_120", lambda: P.laplacian_2d(120), 64, 64, 4),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,max_zero,nb,bins", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_hybrid_factor_and_solve(name, make, max_zero, nb, bins):
    p = make()
    s = mfdag.Solver(nb=nb, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    roots = host_roots(s, p=bins)
    s.set_host_subtrees(roots, threads=4)
    log(f"{name}: n={p.n}, {len(roots)} host subtrees")
    assert s.factorize(p.val) == 0
    ms, nfr, flops = s.host_part()
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    errs = front_errors(s, f, st)
    log(f"{name}: host part {nfr} fronts, {flops / s.cost():.3f} of the flops, {ms:.1f} ms; "
        f"worst front error {errs.max():.2e}")
    assert errs.max() <= 1e-12
    ref = mfdag.Solver(nb=nb, graph=True)
    ref.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    assert ref.factorize(p.val) == 0
    for nrhs in (1, 3):
        xt, b = P.rhs(p, nrhs, seed=20 + nrhs)
        x = s.solve(b.copy())
        x0 = ref.solve(b.copy())
        e0 = np.max(np.abs(x - x0)) / np.max(np.abs(x0))
        eta = P.backward_error(p, x, b)
        log(f"{name} nrhs={nrhs}: vs GPU-only {e0:.2e}, eta {eta:.2e}")
        assert e0 <= 1e-12
        assert eta <= 1e-14
    ref.close()
    s.close()


@pytest.mark.gpu
def test_hybrid_device_values_refactorization_and_forms():
    import ctypes

    p = P.weighted_laplacian_3d(16, seed=5)
    s = mfdag.Solver(nb=32)
    s.analyze(p.ccs_col, p.ccs_row, 64, p.perm)
    s.set_host_subtrees(host_roots(s, p=4), threads=2)
    with pytest.raises(mfdag.MfdError):
        s.factorize(p.val, precision=mfdag.FP32_IR)
    # the values in device memory (hipMalloc of the runtime libmfdag.so uses)
    hip = ctypes.CDLL("libamdhip64.so.7")
    d_a = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_a), ctypes.c_size_t(p.val.nbytes)) == 0
    assert hip.hipMemcpy(d_a, p.val.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(p.val.nbytes), 1) == 0
    xt, b = P.rhs(p, 2, seed=4)
    for scale in (1.0, 2.0):  # two factorizations on one handle
        if scale == 1.0:
            assert s.factorize_device(d_a.value) == 0
        else:
            assert s.factorize(p.val * scale) == 0
        f, st, info = oracle_for(s, p.val * scale)
        errs = front_errors(s, f, st)
        x = s.solve(b.copy())
        r = np.max(np.abs(x * scale - xt)) / np.max(np.abs(xt))
        log(f"scale {scale}: worst front error {errs.max():.2e}, x error {r:.2e}")
        assert errs.max() <= 1e-12
        assert r <= 1e-9
    hip.hipFree(d_a)
    s.set_host_subtrees([])
    assert s.factorize(p.val, precision=mfdag.FP32_IR) == 0
    s.close()
