"""The 16-bit forms (MFD_FP16_IR, MFD_BF16_IR): FP32 fronts of the
diagonally scaled matrix S A S whose update and Schur tiles multiply
16-bit-rounded operands on the 16-bit MFMA, refined in FP64.  Checked: the
scaled fronts against the oracle's FP64 fronts of S A S (within the 16-bit
rounding), x against the oracle's x and x_true, eta <= 1e-14 per right-hand
side at the full 40^3 three-unknown configuration, the fallback to FP32
fronts and FP64 when refinement stalls, and graph replays against eager
runs."""
import sys

import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from parity import front_errors, oracle_for

pytestmark = pytest.mark.gpu
FORMS = [mfdag.FP16_IR, mfdag.BF16_IR]
# relative front error of a factor computed with 16-bit operands in its
# updates (unit roundoff 2^-11 for FP16, 2^-8 for BF16), allowing for growth
FRONT_TOL = {mfdag.FP16_IR: 2e-2, mfdag.BF16_IR: 1.5e-1}


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scaling(p):
    """The library's diagonal scaling: s_j = 2^-floor(E/2), A(j,j) = m 2^E."""
    d = p.scipy().diagonal()
    _, E = np.frexp(d)
    return np.ldexp(1.0, -(E // 2))


@pytest.mark.parametrize("form", FORMS)
def test_scaled_fronts_and_x_against_oracle(form):
    """3-unknown block stencil 8^3: every front within the 16-bit tolerance
    of the oracle's FP64 front of S A S, and the refined x against the
    oracle's x (FP64, unscaled) within the condition-aware bound."""
    p = P.block_stencil_3dof(8, seed=7)
    s = mfdag.Solver(nb=24)
    s.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    assert s.factorize(p.val, precision=form) == 0
    assert s.stats()["precision"] == form
    st = s.structure()
    sc = scaling(p)
    A = p.scipy().tocoo()
    # the scaled values in the caller's CCS order (lower triangle)
    rows = p.ccs_row - 1
    cols = np.repeat(np.arange(p.n), np.diff(p.ccs_col))
    val_s = p.val * sc[rows] * sc[cols]
    f, _, info = oracle_for(s, val_s)
    assert info == 0
    errs = front_errors(s, f, st)
    log(f"form {form}: worst scaled front rel err {errs.max():.2e}")
    assert errs.max() <= FRONT_TOL[form]
    assert errs.max() > 1e-7  # the updates did run in 16 bits
    xt, b = P.rhs(p, 2, seed=3)
    x = s.solve(b)
    stats = s.stats()
    fo, st0, info = oracle_for(s, p.val)
    assert info == 0
    perm = st0["perm"] - 1
    for k in range(2):
        xo = np.empty(p.n)
        xo[perm] = fo.solve(b[perm, k].copy())
        eta = P.backward_error(p, x[:, k], b[:, k])
        rel = np.abs(x[:, k] - xo).max() / np.abs(xo).max()
        log(f"form {form} rhs {k}: iters {stats['refine_iters']}, eta {eta:.2e}, x vs oracle {rel:.2e}")
        assert eta <= 1e-14
        assert rel <= 1e-11
        assert np.abs(x[:, k] - xt[:, k]).max() <= 1e-11 * np.abs(xt[:, k]).max()
    assert stats["precision"] == form and stats["refine_iters"] >= 1
    s.close()


@pytest.mark.parametrize("form", FORMS)
def test_config_block_stencil_3dof_40(form):
    """40^3 grid, 3 unknowns per node (n = 192,000), 27-point block stencil,
    2 right-hand sides: the 16-bit form refines to eta <= 1e-14 per
    right-hand side without leaving its form, and the library's per-rhs
    report agrees with the host."""
    p = P.block_stencil_3dof(40, seed=2024)
    s = mfdag.Solver(nb=64, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 16384, p.perm)
    xt, b = P.rhs(p, 2, seed=2024)
    assert s.factorize(p.val, precision=form) == 0
    x = s.solve(b)
    st = s.stats()
    rep = s.solve_errors(2)
    etas = [P.backward_error(p, x[:, k], b[:, k]) for k in range(2)]
    log(f"blk3_40 form {form}: precision {st['precision']}, iters {st['refine_iters']}, eta {etas}, "
        f"reported {rep['eta']}, ferr est {rep['ferr']}, factor {st['factor_ms']:.2f} ms")
    assert max(etas) <= 1e-14 and rep["eta"].max() <= 1e-14
    assert st["precision"] == form and st["refine_iters"] >= 1
    assert np.abs(x - xt).max() <= 1e-10 * np.abs(xt).max()
    s.close()


def test_fp16_falls_back_on_an_ill_conditioned_matrix():
    """A weighted 3D Laplacian with a 1e-9 shift (condition ~1e12): the FP16
    form cannot refine it, the solve moves to FP32 fronts and then FP64
    (refine_iters negative) and still meets eta <= 1e-14."""
    p = P.weighted_laplacian_3d(12, seed=5, shift=1e-9)
    s = mfdag.Solver(nb=32)
    s.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
    rc = s.factorize(p.val, precision=mfdag.FP16_IR)
    assert rc == 0
    xt, b = P.rhs(p, 1, seed=2)
    x = s.solve(b[:, 0])
    st = s.stats()
    eta = P.backward_error(p, x, b[:, 0])
    log(f"ill-conditioned FP16: final precision {st['precision']}, iters {st['refine_iters']}, eta {eta:.2e}")
    assert eta <= 1e-14
    assert st["precision"] != mfdag.FP16_IR and st["refine_iters"] < 0
    s.close()


@pytest.mark.parametrize("form", FORMS)
def test_lowp_graph_replay_matches_eager(form):
    p = P.weighted_laplacian_3d(14, seed=8, lo=0.5, hi=2.0, shift=1.0)
    xt, b = P.rhs(p, 1, seed=3)
    out = []
    for graph in (False, True):
        s = mfdag.Solver(nb=32, graph=graph)
        s.analyze(p.ccs_col, p.ccs_row, 16, p.perm)
        res = []
        for rep in range(2):
            assert s.factorize(p.val * (1.0 + rep), precision=form) == 0
            res.append((np.concatenate([np.tril(s.front(k)).ravel() for k in range(1, s.num_node + 1)]),
                        s.solve(b[:, 0])))
        out.append(res)
        s.close()
    for (fe, xe), (fg, xg) in zip(*out):
        assert np.array_equal(fe, fg)
        assert np.array_equal(xe, xg)
