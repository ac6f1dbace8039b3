"""The batching form (mfd_analyze_batch): independent systems analysed one
by one and joined into one elimination forest, factored and solved together
by the level schedules.

CPU: the joined structure is each system's own analysis (the same routine
mfd_analyze runs, pinned against the oracle elsewhere) with its columns,
nodes and entries shifted by the systems before it; cost and nonzero add up;
the error paths.
GPU: a batch of anisotropic Laplacians (the batching workload's generator)
in FP64 against x_true and the oracle's solution per system, and in the
FP32 + refinement form; the backward error of every system <= 1e-14."""
import numpy as np
import pytest

import mfdag
from mfdag import problems as P


def small_batch(nsys=6, seed=3, lo=5, hi=14):
    return P.batch_anisotropic(nsys, seed=seed, lo=lo, hi=hi, leaf=8)


def analyze_each(systems, nb, max_zero, with_perm):
    out = []
    for p in systems:
        s = mfdag.Solver(nb=nb, host_only=True)
        s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm if with_perm else None)
        out.append((s.structure(), s.cost(), s.nonzero(), s.num_node, s.n))
        s.close()
    return out


@pytest.mark.parametrize("with_perm", [True, False])
@pytest.mark.parametrize("max_zero", [0, 16])
def test_batch_structure_is_the_systems_joined(with_perm, max_zero):
    systems = small_batch()
    each = analyze_each(systems, 4, max_zero, with_perm)
    s = mfdag.Solver(nb=4, host_only=True)
    s.analyze_batch([(p.ccs_col, p.ccs_row) for p in systems], max_zero,
                    [p.perm for p in systems] if with_perm else None)
    st = s.structure()
    assert s.cost() == sum(e[1] for e in each)
    assert s.nonzero() == sum(e[2] for e in each)
    assert s.num_node == sum(e[3] for e in each)
    assert s.n == sum(e[4] for e in each)
    col = node = ent = a_off = 0
    for (sk, _, _, m, n), p in zip(each, systems):
        sl = slice(node, node + m)
        assert np.array_equal(st["supernode_size"][sl], sk["supernode_size"])
        assert np.array_equal(st["work_size"][sl], sk["work_size"])
        assert np.array_equal(st["parent"][sl], np.where(sk["parent"] > 0, sk["parent"] + node, 0))
        assert np.array_equal(st["perm"][col:col + n], sk["perm"] + col)
        nnz = len(sk["ccs_perm"])
        assert np.array_equal(st["ccs_perm"][a_off:a_off + nnz], sk["ccs_perm"] + ent)
        assert np.array_equal(st["supernodal_index"][a_off:a_off + nnz], sk["supernodal_index"])
        assert np.array_equal(st["a_colptr"][col:col + n], sk["a_colptr"][:n] + a_off)
        col += n
        node += m
        ent += p.nnz
        a_off += nnz
    # one root per system
    assert int(np.sum(st["parent"] == 0)) == len(systems)
    s.close()


def test_reducible_matrix_needs_the_batch_entry():
    systems = small_batch(3)
    bp = P.batch_problem(systems)
    s = mfdag.Solver(nb=4, host_only=True)
    with pytest.raises(mfdag.MfdError) as e:
        s.analyze(bp.ccs_col, bp.ccs_row, 0, None)
    assert e.value.code == -5
    # one malformed system (a column pointer array not starting at 1) fails the batch
    bad = np.asarray(systems[1].ccs_col).copy()
    bad[0] = 2
    with pytest.raises(mfdag.MfdError) as e:
        s.analyze_batch([(systems[0].ccs_col, systems[0].ccs_row), (bad, systems[1].ccs_row)], 0)
    assert e.value.code == -1
    with pytest.raises(mfdag.MfdError):
        s.analyze_batch([], 0)
    s.close()


def batch_rhs(systems, nrhs, seed):
    xs, bs = [], []
    for k, p in enumerate(systems):
        x, b = P.rhs(p, nrhs, seed=seed + k)
        xs.append(x)
        bs.append(b)
    return np.concatenate(xs), np.concatenate(bs)


@pytest.mark.gpu
def test_batch_fp64_against_oracle_and_x_true():
    from parity import front_errors, oracle_for
    systems = P.batch_anisotropic(40, seed=11, lo=16, hi=48)
    bp = P.batch_problem(systems)
    s = mfdag.Solver(nb=64, graph=True)
    s.analyze_batch([(p.ccs_col, p.ccs_row) for p in systems], 64, [p.perm for p in systems])
    assert s.factorize(bp.val) == 0
    f, st, info = oracle_for(s, bp.val)
    assert info == 0
    errs = front_errors(s, f, st)
    assert errs.max() <= 1e-12, (errs.max(), int(errs.argmax()) + 1)
    xt, b = batch_rhs(systems, 2, 5)
    x = s.solve(b.copy())
    perm = st["perm"] - 1
    xo = np.empty(bp.n)
    xo[perm] = f.solve(b[perm, 0].copy())
    assert np.max(np.abs(x[:, 0] - xo)) <= 1e-12 * np.max(np.abs(xo))
    off = 0
    for p in systems:
        xk = x[off:off + p.n]
        assert np.max(np.abs(xk - xt[off:off + p.n])) <= 1e-10
        assert P.backward_error(p, xk, b[off:off + p.n]) <= 1e-14
        off += p.n
    s.close()


@pytest.mark.gpu
def test_batch_fp32_refined():
    systems = P.batch_anisotropic(24, seed=12, lo=16, hi=64)
    bp = P.batch_problem(systems)
    s = mfdag.Solver(nb=64, graph=True)
    s.analyze_batch([(p.ccs_col, p.ccs_row) for p in systems], 64, [p.perm for p in systems])
    xt, b = batch_rhs(systems, 3, 7)
    s.factorize(bp.val, mfdag.FP32_IR)
    x = s.solve(b.copy())
    st = s.stats()
    assert st["refine_iters"] >= 0
    off = 0
    for p in systems:
        assert P.backward_error(p, x[off:off + p.n], b[off:off + p.n]) <= 1e-14
        off += p.n
    s.close()


@pytest.mark.gpu
def test_batch_graph_replays_at_scale():
    """A batch the size of the benchmark's fronts (L up to 160, max_zero
    4096): the captured schedule (main, side and fill streams) replayed three
    times, with eager factorizations and solves between the replays, gives
    the eager results bit for bit, and every system its backward-error
    bound."""
    systems = P.batch_anisotropic(160, seed=13)
    bp = P.batch_problem(systems)
    pairs = [(p.ccs_col, p.ccs_row) for p in systems]
    perms = [p.perm for p in systems]
    eager = mfdag.Solver(nb=64, graph=False)
    graph = mfdag.Solver(nb=64, graph=True)
    for s in (eager, graph):
        s.analyze_batch(pairs, 4096, perms)
    xt, b = batch_rhs(systems, 2, 9)
    assert eager.factorize(bp.val) == 0
    x_ref = eager.solve(b.copy())
    for rep in range(3):
        assert graph.factorize(bp.val) == 0
        x = graph.solve(b.copy())
        assert np.array_equal(x, x_ref), rep
        assert eager.factorize(bp.val) == 0
        assert np.array_equal(eager.solve(b.copy()), x_ref), rep
    off = 0
    for p in systems:
        assert P.backward_error(p, x_ref[off:off + p.n], b[off:off + p.n]) <= 1e-14
        off += p.n
    eager.close()
    graph.close()
