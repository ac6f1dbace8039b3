"""The benchmark configurations at their full sizes, with their results
checked: the 64^3 weighted 7-point Laplacian with 4 right-hand sides (the
headline workload) and the 512-system batch, each in the FP64 and FP32+IR
forms.  The oracle does not reach these sizes in seconds, so the checks are
size-independent: backward error <= 1e-14 per right-hand side (and per
system), the forward error against x_true within a bound set by the
condition number (estimated here by power iterations with the library's own
solves), the two forms' solutions against each other, and the library's own
per-right-hand-side error report against the host's computation."""
import sys

import numpy as np
import pytest

import mfdag
from mfdag import problems as P

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def eta_per_rhs(p, x, b):
    return np.array([P.backward_error(p, x[:, k], b[:, k]) for k in range(x.shape[1])])


def condition_estimate(p, s, iters=30, seed=0):
    """kappa_2(A) = lambda_max / lambda_min of the SPD matrix: power iteration
    with A (scipy) and inverse power iteration with the factor s holds."""
    A = p.scipy()
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(p.n)
    for _ in range(iters):
        v = A @ v
        v /= np.linalg.norm(v)
    lmax = v @ (A @ v)
    u = rng.standard_normal(p.n)
    for _ in range(iters):
        u = s.solve(u)
        u /= np.linalg.norm(u)
    lmin = u @ (A @ u)
    return lmax / lmin


@pytest.fixture(scope="module")
def headline():
    p = P.weighted_laplacian_3d(64, seed=2024)
    xt, b = P.rhs(p, 4, seed=2024)
    return p, xt, b


def test_headline_64cube_both_forms(headline):
    """64^3 weighted 7-point Laplacian (n = 262,144), 4 right-hand sides,
    max_zero 16384, nb 64, captured graphs: both forms reach eta <= 1e-14 on
    every right-hand side, x lies within 4 kappa eta of x_true, the FP32+IR x
    within the same bound of the FP64 x, and mfd_get_solve_errors reports
    the same per-right-hand-side backward errors the host computes."""
    p, xt, b = headline
    s = mfdag.Solver(nb=64, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 16384, p.perm)
    log("64^3: fronts", s.num_node, "cost", s.cost())
    assert s.factorize(p.val) == 0
    x64 = s.solve(b)
    eta64 = eta_per_rhs(p, x64, b)
    rep = s.solve_errors(4)
    kappa = condition_estimate(p, s)
    log(f"64^3 FP64: eta {eta64}, reported {rep['eta']}, ferr est {rep['ferr']}, kappa ~ {kappa:.2e}")
    assert np.all(eta64 <= 1e-14)
    assert np.all(rep["eta"] <= 1e-14)
    assert np.all(rep["eta"] <= 4 * eta64 + 1e-17) and np.all(eta64 <= 4 * rep["eta"] + 1e-17)
    tol = max(1e-12, 4.0 * kappa * eta64.max())
    fwd64 = np.abs(x64 - xt).max(axis=0) / np.abs(xt).max(axis=0)
    log(f"64^3 FP64: forward error vs x_true {fwd64} (bound {tol:.2e})")
    assert np.all(fwd64 <= tol)
    # the estimate is of the same order as the error against x_true
    assert np.all(rep["ferr"] <= 100 * fwd64 + 1e-15) and np.all(fwd64 <= 100 * rep["ferr"] + 1e-15)

    s.factorize(p.val, precision=mfdag.FP32_IR)
    x32 = s.solve(b)
    st = s.stats()
    eta32 = eta_per_rhs(p, x32, b)
    rep32 = s.solve_errors(4)
    log(f"64^3 FP32+IR: iters {st['refine_iters']}, precision {st['precision']}, eta {eta32}, "
        f"reported {rep32['eta']}, ferr est {rep32['ferr']}")
    assert st["refine_iters"] >= 1
    assert np.all(eta32 <= 1e-14) and np.all(rep32["eta"] <= 1e-14)
    fwd32 = np.abs(x32 - xt).max(axis=0) / np.abs(xt).max(axis=0)
    tol32 = max(1e-12, 4.0 * kappa * max(eta32.max(), eta64.max()))
    d = np.abs(x32 - x64).max(axis=0) / np.abs(x64).max(axis=0)
    log(f"64^3 FP32+IR: forward error vs x_true {fwd32}, vs FP64 x {d} (bound {tol32:.2e})")
    assert np.all(fwd32 <= tol32) and np.all(d <= tol32)
    s.close()


@pytest.mark.parametrize("precision", [mfdag.FP64, mfdag.FP32_IR])
def test_batch_512_systems(precision):
    """The batching form at its full size: 512 anisotropic 2D Laplacians
    (L ~ U{16..160}), max_zero 4096, one right-hand side each, captured
    graphs; every system meets eta <= 1e-14 and x_true within 1e-8 (the
    systems' kappa is at most ~1e5)."""
    systems = P.batch_anisotropic(512, seed=2024)
    bp = P.batch_problem(systems)
    xs, bs = zip(*[P.rhs(p, 1, seed=2024 + k) for k, p in enumerate(systems)])
    xt, b = np.concatenate(xs), np.concatenate(bs)
    s = mfdag.Solver(nb=64, graph=True)
    s.analyze_batch([(p.ccs_col, p.ccs_row) for p in systems], 4096, [p.perm for p in systems])
    log("batch 512: n", bp.n, "fronts", s.num_node, "precision", precision)
    assert s.factorize(bp.val, precision=precision) == 0
    x = s.solve(b.copy())
    st = s.stats()
    worst_eta, worst_x, off = 0.0, 0.0, 0
    for p in systems:
        xk, bk = x[off:off + p.n], b[off:off + p.n]
        worst_eta = max(worst_eta, P.backward_error(p, xk, bk))
        worst_x = max(worst_x, np.abs(xk - xt[off:off + p.n]).max() / np.abs(xt[off:off + p.n]).max())
        off += p.n
    rep = s.solve_errors(1, forward=False)
    log(f"batch 512 precision {precision}: worst eta {worst_eta:.2e}, worst x err {worst_x:.2e}, "
        f"iters {st['refine_iters']}, reported eta {rep['eta'][0]:.2e}")
    assert worst_eta <= 1e-14
    assert worst_x <= 1e-8
    # the whole block-diagonal system's eta is bounded by its worst system's
    assert rep["eta"][0] <= 1e-14
    s.close()
