"""GPU parity: the HIP numeric phase against the oracle (the C restatement of
the reference's seq_factorize / seq_forward / seq_backward) and against the
reference's own known answers.  Tolerances: factor fronts and solutions
within 1e-12 relative (normwise per front / per solution), backward error
<= 1e-14; integer-valued known answers exactly."""
import ctypes
import sys

import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from oracle import oracle as O

from parity import front_errors, oracle_for, tile_mask_compare
from test_oracle_kats import check_three_front_factors

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def three_front(kats, nb, graph=False):
    t = kats["three_front"]
    s = mfdag.Solver(nb=nb, graph=graph)
    lens = np.array(t["supernodal_index_lengths"], np.int64)
    colptr = np.concatenate([[0], np.cumsum(lens)])
    s.set_structure(t["supernode_size"], t["work_size"], t["parent"], t["local_index"], colptr,
                    t["supernodal_index"])
    return s, np.array(t["val"], float)


@pytest.mark.parametrize("nb", [2, 3])
def test_three_front_kat_exact(kats, nb):
    """seq_factorize_test.f90 (13 x 13, integer factors): exact tiles."""
    log("three_front nb", nb)
    s, val = three_front(kats, nb)
    assert s.factorize(val) == 0
    sn, _ = s.export_tiles()
    f = O.OracleFactor(kats["three_front"]["supernode_size"], kats["three_front"]["work_size"],
                       kats["three_front"]["parent"], nb, kats["three_front"]["local_index"])

    def tile(node, i, j):
        off, r, c = f.tile(1, node, i, j)
        return sn[off:off + r * c]
    check_three_front_factors(kats, nb, tile)


def test_three_front_solve(kats):
    """Solve with the factor of the 13 x 13 problem for x = 1 (the known
    answer of seq_forward_test / seq_backward_test)."""
    s, val = three_front(kats, 2)
    s.factorize(val)
    st = s.structure()
    f, _, _ = oracle_for(s, val)
    b = np.ones(13)
    xo = f.solve(b.copy())
    x = s.solve(b, permuted=True)
    assert np.allclose(x, xo, rtol=1e-13, atol=1e-13)
    # x = A^{-1} (A 1): build A 1 from the oracle's dense front data
    A = assemble_dense_from_parts(kats)
    x1 = s.solve(A @ np.ones(13), permuted=True)
    assert np.allclose(x1, 1.0, rtol=0, atol=1e-12)


def assemble_dense_from_parts(kats):
    t = kats["three_front"]
    c, w, par, li = t["supernode_size"], t["work_size"], t["parent"], t["local_index"]
    first = np.concatenate([[1], 1 + np.cumsum(c)])
    lptr = np.concatenate([[0], np.cumsum(w)])
    rows_of = {}
    for s in range(len(c) - 1, -1, -1):
        p = par[s]
        rows = []
        for k in range(w[s]):
            pos = li[lptr[s] + k]
            rows.append(first[p - 1] + pos - 1 if pos <= c[p - 1] else rows_of[p - 1][pos - c[p - 1] - 1])
        rows_of[s] = rows
    n = sum(c)
    A = np.zeros((n, n))
    lens = t["supernodal_index_lengths"]
    k = 0
    col = 0
    for s in range(len(c)):
        for j in range(c[s]):
            for _ in range(lens[col]):
                r = t["supernodal_index"][k]
                g = first[s] + r - 1 if r <= c[s] else rows_of[s][r - c[s] - 1]
                A[g - 1, col] = t["val"][k]
                A[col, g - 1] = t["val"][k]
                k += 1
            col += 1
    return A


CASES = [
    ("lap2d_20", lambda: P.laplacian_2d(20), 0, 8),
    ("lap2d_100", lambda: P.laplacian_2d(100), 16, 64),
    ("wlap3d_8", lambda: P.weighted_laplacian_3d(8, seed=3), 0, 16),
    ("wlap3d_12x10x9", lambda: P.weighted_laplacian_3d(12, 10, 9, seed=4), 32, 5),
    ("st27_7", lambda: P.stencil27_3d(7, seed=5), 8, 7),
    ("wlap3d_16", lambda: P.weighted_laplacian_3d(16, seed=6), 64, 32),
    ("blk3_8", lambda: P.block_stencil_3dof(8, seed=7), 16, 24),
]


@pytest.mark.parametrize("name,make,max_zero,nb", CASES)
def test_factor_and_solve_vs_oracle(name, make, max_zero, nb):
    log("case", name, "max_zero", max_zero, "nb", nb)
    p = make()
    s = mfdag.Solver(nb=nb)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    assert s.factorize(p.val) == 0
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    errs = front_errors(s, f, st)
    assert errs.max() <= 1e-12, (name, errs.max(), int(errs.argmax()) + 1)
    worst, cnt = tile_mask_compare(s, f)
    assert cnt > 0 and worst <= 1e-12, (name, worst)
    xt, b = P.rhs(p, 1, seed=1)
    x = s.solve(b[:, 0])
    # oracle solve in the permuted numbering
    perm = st["perm"] - 1
    xo = np.empty(p.n)
    xo[perm] = f.solve(b[perm, 0].copy())
    eta, eta_o = P.backward_error(p, x, b[:, 0]), P.backward_error(p, xo, b[:, 0])
    assert eta <= 1e-14 and eta_o <= 1e-14
    # Two backward-stable solutions differ by up to ~ 2 kappa eta; the
    # bound is 1e-12 unless the condition number makes that unreachable.
    tol = max(1e-12, 4.0 * condition_number(p) * max(eta, eta_o))
    rel = np.linalg.norm(x - xo, np.inf) / np.linalg.norm(xo, np.inf)
    log(f"{name}: x rel diff {rel:.2e} (tol {tol:.2e}), eta {eta:.2e}")
    assert rel <= tol, (name, rel, tol)


def condition_number(p):
    """2-norm condition number of the SPD matrix (extreme eigenvalues)."""
    from scipy.sparse.linalg import eigsh
    A = p.scipy()
    if p.n <= 2000:
        ev = np.linalg.eigvalsh(A.toarray())
        return ev[-1] / ev[0]
    hi = eigsh(A, k=1, which="LA", return_eigenvectors=False)[0]
    lo = eigsh(A, k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    return hi / lo


def test_multiple_rhs_and_device_entry_points():
    p = P.weighted_laplacian_3d(10, seed=7)
    s = mfdag.Solver(nb=16)
    s.analyze(p.ccs_col, p.ccs_row, 8, p.perm)
    s.factorize(p.val)
    xt, b = P.rhs(p, 4, seed=2)
    x = s.solve(b)
    for k in range(4):
        xk = s.solve(b[:, k])
        assert np.array_equal(xk, x[:, k])
    assert P.backward_error(p, x, b) <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("rule", [0, 1])
def test_level_rules_against_oracle(rule):
    """Both level rules (fronts one level above their highest child, or one
    below their parent) give the oracle's fronts and the same solution."""
    p = P.weighted_laplacian_3d(20, seed=3)
    s = mfdag.Solver(nb=32, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 64, p.perm)
    s.level_rule = rule
    assert s.level_rule == rule
    assert s.factorize(p.val) == 0
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    errs = front_errors(s, f, st)
    xt, b = P.rhs(p, 2, seed=9)
    x = s.solve(b.copy())
    eta = P.backward_error(p, x, b)
    print(f"rule {rule}: worst front error {errs.max():.2e}, eta {eta:.2e}", flush=True)
    assert errs.max() <= 1e-12
    assert eta <= 1e-14
    s.close()


@pytest.mark.parametrize("precision", [mfdag.FP64, mfdag.FP32_IR])
def test_more_right_hand_sides_than_one_group(precision):
    """11 right-hand sides: the solve kernels take them 8 at a time, so the
    second group is partial; each column equals its own single solve bit for
    bit (FP64; the refined form is checked by its backward error)."""
    p = P.weighted_laplacian_3d(11, seed=5)
    s = mfdag.Solver(nb=32, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    s.factorize(p.val, precision=precision)
    xt, b = P.rhs(p, 11, seed=6)
    x = s.solve(b)
    if precision == mfdag.FP64:
        for k in (0, 7, 8, 10):
            assert np.array_equal(s.solve(b[:, k]), x[:, k]), k
    assert P.backward_error(p, x, b) <= 1e-14
    assert np.max(np.abs(x - xt)) <= 1e-8 * np.max(np.abs(xt))
    # ldb > n through the C-ABI: the same solution, rows past n untouched
    ld = p.n + 5
    buf = np.full((ld, 11), 7.0, order="F")
    buf[:p.n] = b
    rc = mfdag.lib().mfd_solve(s.h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 11, ld)
    assert rc == 0
    assert np.array_equal(buf[:p.n], x)
    assert np.all(buf[p.n:] == 7.0)
    s.close()


def test_graph_replay_matches_eager():
    """hipGraph replays (several, with eager work in between) give the
    bitwise eager result."""
    p = P.weighted_laplacian_3d(14, seed=8)
    eager = mfdag.Solver(nb=32)
    eager.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    eager.factorize(p.val)
    fronts = [eager.front(k) for k in range(1, eager.num_node + 1)]
    g = mfdag.Solver(nb=32, graph=True)
    g.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    xt, b = P.rhs(p, 1, seed=3)
    for rep in range(3):
        g.factorize(p.val * (1.0 + rep))
        eager.factorize(p.val * (1.0 + rep))
        fronts = [eager.front(k) for k in range(1, eager.num_node + 1)]
        for k in range(1, g.num_node + 1):
            assert np.array_equal(np.tril(g.front(k)), np.tril(fronts[k - 1])), (rep, k)
        assert np.array_equal(g.solve(b[:, 0]), eager.solve(b[:, 0]))


def test_fp32_refinement_reaches_fp64_accuracy():
    p = P.weighted_laplacian_3d(16, seed=9)
    s = mfdag.Solver(nb=64)
    s.analyze(p.ccs_col, p.ccs_row, 32, p.perm)
    s.factorize(p.val, precision=mfdag.FP32_IR)
    xt, b = P.rhs(p, 2, seed=4)
    x = s.solve(b)
    st = s.stats()
    log("fp32 ir iters", st["refine_iters"], "eta", st["backward_error"])
    assert st["refine_iters"] >= 1
    assert st["backward_error"] <= 1e-14
    assert P.backward_error(p, x, b) <= 1e-14
    # ||b - A x||_2 / ||x||_2 (example.f90:147-180): the residual is rounding
    # noise, so the host's and the device's sums agree in magnitude only.
    r = b - p.scipy() @ x
    host = max(np.linalg.norm(r[:, k]) / np.linalg.norm(x[:, k]) for k in range(2))
    log("fp32 ir res2", st["refine_residual"], "host", host)
    assert 0 < st["refine_residual"] <= 1e-12
    assert host / 10 <= st["refine_residual"] <= host * 10


@pytest.mark.parametrize("graph", [False, True])
def test_refactorize_same_structure_new_values(graph):
    """Values-only refactorization on one analysis (the reference calls
    hppllt_factorize again with the structure of hppllt_analyze,
    hppllt.f90:123-150): new values, then the first values again, then the
    FP32+IR form and FP64 again; each factor against the oracle's for its own
    values, each x to the backward-error bound."""
    log("refactorize graph", graph)
    p1 = P.weighted_laplacian_3d(12, seed=3)
    p2 = P.weighted_laplacian_3d(12, seed=4)
    assert np.array_equal(p1.ccs_row, p2.ccs_row) and not np.array_equal(p1.val, p2.val)
    s = mfdag.Solver(nb=32, graph=graph)
    s.analyze(p1.ccs_col, p1.ccs_row, 16, p1.perm)
    for prob, precision in ((p1, mfdag.FP64), (p2, mfdag.FP64), (p1, mfdag.FP64), (p2, mfdag.FP32_IR),
                            (p2, mfdag.FP64)):
        s.factorize(prob.val, precision=precision)
        if precision == mfdag.FP64:
            f, st, info = oracle_for(s, prob.val)
            assert info == 0
            assert front_errors(s, f, st).max() <= 1e-12
        xt, b = P.rhs(prob, 2, seed=7)
        x = s.solve(b)
        assert P.backward_error(prob, x, b) <= 1e-14
    s.close()


def test_fp32_ir_falls_back_to_fp64_on_a_nearly_singular_matrix():
    """A weighted graph Laplacian with a 1e-9 shift (condition ~1e12): FP32
    cannot factor it usefully (a pivot of the shift's size drowns in FP32
    rounding of the ~1e3 entries), so the precision form either refactors in
    FP64 at once (non-positive FP32 pivot) or after refinement stalls; either
    way the solve reaches the FP64 bound."""
    p = P.weighted_laplacian_3d(12, seed=3, shift=1e-9)
    s = mfdag.Solver(nb=32)
    s.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    assert s.factorize(p.val, precision=mfdag.FP32_IR) == 0
    xt, b = P.rhs(p, 1, seed=8)
    x = s.solve(b)
    st = s.stats()
    log("nearly singular: precision", st["precision"], "refine_iters", st["refine_iters"])
    assert st["precision"] == mfdag.FP64 or st["refine_iters"] < 0
    assert P.backward_error(p, x, b) <= 1e-14
    s.close()


def test_fp32_pivot_failure_refactors_in_fp64():
    """An SPD matrix whose FP32 factorization meets a zero pivot: a 10 x 10
    Laplacian plus a pair of unknowns with A = [[1, -(1 - 2^-30)], [., 1]]
    (eigenvalue 2^-30; its off-diagonal rounds to -1 in FP32, leaving the
    pair's second pivot exactly 0 there) tied to the grid by a 1e-8
    coupling.  FP64 factors it; the precision form must refactor in FP64
    instead of reporting a non-positive-definite matrix."""
    g = P.laplacian_2d(10, leaf=1000)
    n = g.n + 2
    cols = np.repeat(np.arange(g.n), np.diff(g.ccs_col))
    rows = g.ccs_row - 1
    rows = np.concatenate([rows, [g.n, g.n + 1, g.n + 1, g.n]])
    cols = np.concatenate([cols, [g.n, g.n + 1, g.n, 5]])
    vals = np.concatenate([g.val, [1.0, 1.0, -(1.0 - 2.0 ** -30), -1e-8]])
    cc, rr, vv = P.lower_ccs(n, rows, cols, vals)
    s = mfdag.Solver(nb=16)
    s.analyze(cc, rr, 0, None)
    assert s.factorize(vv, precision=mfdag.FP32_IR) == 0
    assert s.stats()["precision"] == mfdag.FP64
    prob = P.Problem(n, cc, rr, vv)
    xt, b = P.rhs(prob, 1, seed=4)
    x = s.solve(b)
    assert P.backward_error(prob, x, b) <= 1e-14
    s.close()


def test_reanalysis_on_one_handle():
    """One handle (captured graphs on) analysed again and again: a 3D
    problem, a 2D one of another size, the 3D one again, then a batch; each
    factorization and solve is correct, so nothing captured for an earlier
    analysis is replayed."""
    p3 = P.weighted_laplacian_3d(10, seed=2)
    p2 = P.laplacian_2d(37)
    systems = P.batch_anisotropic(6, seed=3, lo=12, hi=40)
    s = mfdag.Solver(nb=32, graph=True)
    for prob in (p3, p2, p3):
        s.analyze(prob.ccs_col, prob.ccs_row, 16, prob.perm)
        for precision in (mfdag.FP64, mfdag.FP32_IR):
            s.factorize(prob.val, precision=precision)
            xt, b = P.rhs(prob, 2, seed=1)
            assert P.backward_error(prob, s.solve(b), b) <= 1e-14, (prob.name, precision)
    s.analyze_batch([(q.ccs_col, q.ccs_row) for q in systems], 16, [q.perm for q in systems])
    bp = P.batch_problem(systems)
    s.factorize(bp.val)
    xt, b = P.rhs(bp, 1, seed=2)
    assert P.backward_error(bp, s.solve(b), b) <= 1e-14
    s.close()


def test_not_positive_definite_reports_column():
    p = P.laplacian_2d(12)
    val = p.val.copy()
    diag_pos = p.ccs_col[:-1] - 1
    val[diag_pos[50]] = -4.0   # column 51 (caller numbering) loses positivity
    s = mfdag.Solver(nb=8)
    s.analyze(p.ccs_col, p.ccs_row, 0, p.perm)
    with pytest.raises(mfdag.NotPositiveDefinite) as e:
        s.factorize(val)
    f, st, info = oracle_for(s, val)
    assert info > 0
    assert e.value.code >= 1
    with pytest.raises(mfdag.MfdError):
        s.solve(np.ones(p.n))   # no valid factor
    # the precision form reports it too (its FP64 refactorization fails as well)
    with pytest.raises(mfdag.NotPositiveDefinite) as e32:
        s.factorize(val, precision=mfdag.FP32_IR)
    assert e32.value.code == e.value.code


def test_memory_budget_is_an_error(monkeypatch):
    monkeypatch.setenv("MFD_DEVICE_MEM_LIMIT", "100000")
    p = P.weighted_laplacian_3d(12, seed=1)
    s = mfdag.Solver(nb=16)
    with pytest.raises(mfdag.MfdError) as e:
        s.analyze(p.ccs_col, p.ccs_row, 0, p.perm)
    assert e.value.code == -4


@pytest.mark.parametrize("n", [1, 40, 65, 200])
def test_dense_front_matches_lapack(n):
    """A dense matrix is one front (c = n, w = 0): exercises the small-front,
    panel and update kernels against numpy's Cholesky."""
    rng = np.random.default_rng(n)
    m = rng.standard_normal((n, n))
    Ad = m @ m.T + n * np.eye(n)
    rows, cols = np.tril_indices(n)
    cc, rr, vv = P.lower_ccs(n, rows, cols, Ad[rows, cols])
    s = mfdag.Solver(nb=16)
    s.analyze(cc, rr, 0)
    s.factorize(vv)
    assert s.num_node == 1
    F = s.front(1)
    perm = s.structure()["perm"] - 1
    L = np.linalg.cholesky(Ad[np.ix_(perm, perm)])
    assert np.abs(np.tril(F) - L).max() <= 1e-12 * np.abs(L).max()


# ---------------------------------------------------------------- configs
# The benchmark configurations at full size, checked through properties that
# do not depend on size (the oracle is checked on the smaller cases above).

def solve_checks(p, max_zero, nb, precision, nrhs=2, seed=11):
    s = mfdag.Solver(nb=nb, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    assert s.factorize(p.val, precision) == 0
    xt, b = P.rhs(p, nrhs, seed=seed)
    x = s.solve(b)
    eta = P.backward_error(p, x, b)
    st = s.stats()
    return s, xt, x, eta, st


def test_config_lap2d_100_x_true():
    """2D 5-point Laplacian, 100 x 100 (n = 10,000), x_true ~ U(-1,1):
    x within 1e-12 of x_true (kappa ~ 6e3) and eta <= 1e-14."""
    p = P.laplacian_2d(100)
    s, xt, x, eta, st = solve_checks(p, 16, 64, mfdag.FP64, nrhs=1)
    rel = np.linalg.norm(x - xt, np.inf) / np.linalg.norm(xt, np.inf)
    log(f"lap2d_100: x err {rel:.2e}, eta {eta:.2e}")
    assert rel <= 1e-12 and eta <= 1e-14


@pytest.mark.parametrize("precision", [mfdag.FP64, mfdag.FP32_IR])
def test_config_stencil27_48(precision):
    """48^3 27-point stencil (n = 110,592), weights log-uniform in
    [1e-1, 1e1], shift 1e-4: eta <= 1e-14 in both forms."""
    p = P.stencil27_3d(48, seed=2024)
    s, xt, x, eta, st = solve_checks(p, 16384, 64, precision)
    log(f"st27_48 precision {precision}: eta {eta:.2e} (reported {st['backward_error']:.2e}), "
        f"iters {st['refine_iters']}")
    assert eta <= 1e-14


@pytest.mark.parametrize("precision", [mfdag.FP64, mfdag.FP32_IR])
def test_config_block_stencil_3dof_40(precision):
    """40^3 grid with 3 unknowns per node (n = 192,000), 27-point 3 x 3 block
    stencil: eta <= 1e-14 in both forms."""
    p = P.block_stencil_3dof(40, seed=2024)
    s, xt, x, eta, st = solve_checks(p, 16384, 64, precision)
    log(f"blk3_40 precision {precision}: eta {eta:.2e}, iters {st['refine_iters']}")
    assert eta <= 1e-14


@pytest.mark.parametrize("precision", [mfdag.FP64, mfdag.FP32_IR])
def test_solve_errors_per_rhs_and_fp32_ir_x_against_oracle(precision):
    """mfd_get_solve_errors: per right-hand side, the backward error and the
    reference driver's ||b - Ax||_2/||x||_2 equal the host's values (to the
    rounding of the residual), and the forward error estimate is of the order
    of the error against the oracle's FP64 x.  The FP32+IR x agrees with the
    oracle's x within the condition-aware bound of the FP64 parity test.
    Before a solve, after a new factorization, and for another nrhs the call
    is refused."""
    p = P.weighted_laplacian_3d(12, seed=21)
    s = mfdag.Solver(nb=32, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    assert s.factorize(p.val, precision=precision) == 0
    with pytest.raises(mfdag.MfdError):
        s.solve_errors(3)
    xt, b = P.rhs(p, 3, seed=22)
    b[:, 1] *= 1e3  # right-hand sides of different scales
    x = s.solve(b)
    rep = s.solve_errors(3)
    with pytest.raises(mfdag.MfdError):
        s.solve_errors(2)
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    perm = st["perm"] - 1
    A = p.scipy()
    kappa = condition_number(p)
    for k in range(3):
        eta = P.backward_error(p, x[:, k], b[:, k])
        r = b[:, k] - A @ x[:, k]
        res2 = np.linalg.norm(r) / np.linalg.norm(x[:, k])
        xo = np.empty(p.n)
        xo[perm] = f.solve(b[perm, k].copy())
        fwd = np.abs(x[:, k] - xo).max() / np.abs(xo).max()
        tol = max(1e-12, 4.0 * kappa * max(eta, P.backward_error(p, xo, b[:, k])))
        log(f"rhs {k}: eta {eta:.2e} reported {rep['eta'][k]:.2e}; res2 {res2:.2e} reported "
            f"{rep['res2'][k]:.2e}; x vs oracle {fwd:.2e} (tol {tol:.2e}) ferr est {rep['ferr'][k]:.2e}")
        assert eta <= 1e-14 and rep["eta"][k] <= 1e-14
        assert rep["eta"][k] <= 4 * eta + 1e-18 and eta <= 4 * rep["eta"][k] + 1e-18
        assert res2 / 10 <= rep["res2"][k] <= res2 * 10
        assert fwd <= tol
        assert rep["ferr"][k] <= tol
    assert s.factorize(p.val, precision=precision) == 0
    with pytest.raises(mfdag.MfdError):
        s.solve_errors(3)
    s.close()
