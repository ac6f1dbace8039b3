"""Forward and backward solves over 256-column superblocks (k_superinv,
k_fwd_super, k_bwd_super), chosen per front by mfd_set_solve_block.  No
counterpart in the reference (its sweeps are one DTRSV and one DGEMV per
front, seq_solve/seq_forward_m.f90:14-81, seq_solve/seq_backward_m.f90:
14-131); the solution it produces is the reference's.

CPU: the setting's argument checks (no compute without a GPU).
GPU: with a bound low enough that fronts of one partial superblock (with and
without update rows), of several superblocks and of 64-column blocks share
levels, x against the oracle's x (1e-10 relative to max |x|) and against the
64-column path (1e-12), eta <= 1e-14 for 1, 3 and 9 right-hand sides (one
and two groups of eight), and the FP32+IR and BF16+IR forms refined to eta
<= 1e-14 through the same sweep."""
import sys

import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from parity import oracle_for


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------- CPU
def test_solve_block_checked_without_compute():
    p = P.laplacian_2d(20)
    s = mfdag.Solver(nb=16, host_only=True)
    assert s.solve_block == 256
    with pytest.raises(mfdag.MfdError):
        s.solve_block = -1
    s.analyze(p.ccs_col, p.ccs_row, 8, p.perm)
    s.solve_block = 65
    assert s.solve_block == 65
    s.solve_block = 0
    assert s.solve_block == 0
    s.close()


# ------------------------------------------------------------------- GPU
# largest fronts (pivot columns, update rows): blk3_12 (612, 0), (216, 432),
# (108, 468); wlap3d_20 (569, 0), (180, 400); lap2d_120 (120-ish, 0)
CASES = [
    ("blk3_12", lambda: P.block_stencil_3dof(12, seed=7), 64, 32),
    ("wlap3d_20", lambda: P.weighted_laplacian_3d(20, seed=3), 64, 32),
    ("lap2d_120", lambda: P.laplacian_2d(120), 64, 64),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,max_zero,nb", CASES, ids=[c[0] for c in CASES])
def test_superblock_forward_against_oracle(name, make, max_zero, nb):
    p = make()
    s = mfdag.Solver(nb=nb, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    s.solve_block = 65
    log(f"{name}: n={p.n} superblock forward from 65 pivot columns")
    assert s.factorize(p.val) == 0
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    perm = st["perm"] - 1
    ref = mfdag.Solver(nb=nb, graph=True)
    ref.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    ref.solve_block = 0
    assert ref.factorize(p.val) == 0
    for nrhs in (1, 3, 9):
        xt, b = P.rhs(p, nrhs, seed=11 + nrhs)
        x = s.solve(b.copy())
        x0 = ref.solve(b.copy())
        xo = np.empty(p.n)
        xo[perm] = f.solve(b[perm, 0].copy())
        scale = np.max(np.abs(xo))
        e_or = np.max(np.abs(x[:, 0] - xo)) / scale
        e_64 = np.max(np.abs(x - x0)) / np.max(np.abs(x0))
        eta = P.backward_error(p, x, b)
        log(f"{name} nrhs={nrhs}: vs oracle {e_or:.2e}, vs 64-column path {e_64:.2e}, eta {eta:.2e}")
        assert e_or <= 1e-10, (name, nrhs, e_or)
        assert e_64 <= 1e-12, (name, nrhs, e_64)
        assert eta <= 1e-14, (name, nrhs, eta)
    ref.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mfdag.FP32_IR, mfdag.BF16_IR], ids=["fp32_ir", "bf16_ir"])
def test_superblock_forward_refined_forms(precision):
    p = P.block_stencil_3dof(12, seed=7)
    s = mfdag.Solver(nb=32, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 64, p.perm)
    s.solve_block = 65
    assert s.factorize(p.val, precision=precision) == 0
    xt, b = P.rhs(p, 3, seed=5)
    x = s.solve(b.copy())
    eta = P.backward_error(p, x, b)
    err = np.max(np.abs(x - xt)) / np.max(np.abs(xt))
    log(f"precision {precision}: eta {eta:.2e}, forward error {err:.2e}")
    assert eta <= 1e-14, eta
    s.close()
