"""The CPU comparator (bench.py's cpu_baseline): the oracle's OpenMP-task
restatement of the reference's StarPU task graph (orc_factorize_dag) gives
the sequential oracle's factor (orc_factorize, itself pinned by the
reference's known answers) on the same problem, to the rounding that a
different order of the commuting updates allows, with the oracle's own tile
kernels and with the host BLAS, on one and on several threads; the
parallel assembly gives the sequential one's tiles exactly."""
import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from oracle import oracle as O


def setup(p, nb, max_zero):
    s = mfdag.Solver(nb=nb, host_only=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    st = s.structure()
    return st, p.val[st["ccs_perm"] - 1]


def factor(st, vals, nb, threads=None):
    f = O.OracleFactor(st["supernode_size"], st["work_size"], st["parent"], nb, st["local_index"])
    if threads is None:
        f.assemble(st["a_colptr"], st["supernodal_index"], vals)
        return f, f.factorize()
    f.assemble_par(st["a_colptr"], st["supernodal_index"], vals, threads)
    return f, f.factorize_dag(threads)


CASES = [("wlap3d_12", lambda: P.weighted_laplacian_3d(12, seed=2), 16, 64),
         ("st27_8", lambda: P.stencil27_3d(8, seed=3), 24, 4096),
         ("blk3_6", lambda: P.block_stencil_3dof(6, seed=4), 20, 16),
         ("lap2d_30", lambda: P.laplacian_2d(30), 7, 8)]


@pytest.mark.parametrize("blas", [False, True])
@pytest.mark.parametrize("name,make,nb,max_zero", CASES)
def test_dag_matches_sequential_oracle(name, make, nb, max_zero, blas):
    p = make()
    st, vals = setup(p, nb, max_zero)
    O.textbook_kernels()
    ref, info = factor(st, vals, nb)
    assert info == 0
    if blas and O.host_blas() is None:
        pytest.skip("no host BLAS in this image")
    xt, b = P.rhs(p, 1, seed=5)
    perm = st["perm"] - 1
    xref = np.empty(p.n)
    xref[perm] = ref.solve(b[perm, 0].copy())
    try:
        for threads in (1, 4):
            f, info = factor(st, vals, nb, threads)
            assert info == 0
            worst = 0.0
            for k in range(1, len(st["supernode_size"]) + 1):
                a, r = np.tril(f.front(k)), np.tril(ref.front(k))
                worst = max(worst, np.linalg.norm(a - r) / np.linalg.norm(r))
            assert worst <= 1e-13, (name, threads, worst)
            x = np.empty(p.n)
            x[perm] = f.solve(b[perm, 0].copy())
            assert np.abs(x - xref).max() <= 1e-10 * np.abs(xref).max()
            assert P.backward_error(p, x, b[:, 0]) <= 1e-14
    finally:
        O.textbook_kernels()


def test_parallel_assembly_is_the_sequential_one():
    p = P.weighted_laplacian_3d(10, seed=1)
    st, vals = setup(p, 16, 32)
    a = O.OracleFactor(st["supernode_size"], st["work_size"], st["parent"], 16, st["local_index"])
    b = O.OracleFactor(st["supernode_size"], st["work_size"], st["parent"], 16, st["local_index"])
    a.assemble(st["a_colptr"], st["supernodal_index"], vals)
    b.assemble_par(st["a_colptr"], st["supernodal_index"], vals, 4)
    for which in (1, 2, 3):
        assert np.array_equal(a.array(which), b.array(which))


def test_dag_reports_a_non_positive_pivot():
    p = P.weighted_laplacian_3d(8, seed=1)
    val = p.val.copy()
    d = p.ccs_col[:-1] - 1  # the diagonal entries (first in each column)
    val[d[200]] = -1.0
    st, vals = setup(P.Problem(p.n, p.ccs_col, p.ccs_row, val, p.perm), 16, 16)
    _, info_seq = factor(st, vals, 16)
    _, info_dag = factor(st, vals, 16, 4)
    assert info_seq > 0 and info_dag == info_seq
