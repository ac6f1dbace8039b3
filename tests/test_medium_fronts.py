"""Medium fronts: the whole partial factorization of a front by one
workgroup (k_front_wg), chosen per front by mfd_set_front_split /
mfd_set_front_kinds.  No counterpart in the reference (its StarPU tasks are
per tile); the factor it produces is the reference's.

CPU: the settings' argument and state checks (no compute without a GPU).
GPU: every front against the oracle (FP64, <= 1e-12 relative) with all
fronts of order > 64 medium and with medium and large fronts mixed per node
(a medium child adding into a large parent and the reverse), x against the
oracle's x, eta <= 1e-14 in every form, graph replay equal to eager runs, the
per-front clocks."""
import sys

import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from parity import front_errors, oracle_for


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------- CPU
def test_settings_checked_without_compute():
    p = P.laplacian_2d(20)
    s = mfdag.Solver(nb=16, host_only=True)
    with pytest.raises(mfdag.MfdError):
        s.set_front_split(-1)
    with pytest.raises(mfdag.MfdError) as e:  # per-node choices need an analysis
        s.set_front_kinds(np.zeros(3, dtype=np.int32))
    assert e.value.code == mfdag.MFD_ERR_STATE
    s.analyze(p.ccs_col, p.ccs_row, 8, p.perm)
    s.set_front_split(256)
    bad = np.zeros(s.num_node, dtype=np.int32)
    bad[0] = 2
    with pytest.raises(mfdag.MfdError):
        s.set_front_kinds(bad)
    s.set_front_kinds(np.ones(s.num_node, dtype=np.int32))
    s.set_front_kinds(None)
    with pytest.raises(mfdag.MfdError) as e:
        s.front_times()
    assert e.value.code == mfdag.MFD_ERR_NODEVICE
    s.close()


# ------------------------------------------------------------------- GPU
CASES = [
    ("lap2d_100", lambda: P.laplacian_2d(100), 64, 64),
    ("wlap3d_16", lambda: P.weighted_laplacian_3d(16, seed=3), 256, 32),
    ("st27_10", lambda: P.stencil27_3d(10, seed=2), 128, 24),
    ("blk3_8", lambda: P.block_stencil_3dof(8, seed=7), 64, 16),
]


def _check_fp64(s, p, tag):
    assert s.factorize(p.val) == 0
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    errs = front_errors(s, f, st)
    log(f"{tag}: worst front rel err {errs.max():.2e}")
    assert errs.max() <= 1e-12, (tag, errs.max(), int(errs.argmax()) + 1)
    xt, b = P.rhs(p, 2, seed=4)
    x = s.solve(b.copy())
    perm = st["perm"] - 1
    xo = np.empty(p.n)
    xo[perm] = f.solve(b[perm, 0].copy())
    assert np.max(np.abs(x[:, 0] - xo)) <= 1e-10 * np.max(np.abs(xo)), tag
    assert P.backward_error(p, x, b) <= 1e-14, tag


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,max_zero,nb", CASES, ids=[c[0] for c in CASES])
def test_all_medium_against_oracle(name, make, max_zero, nb):
    p = make()
    s = mfdag.Solver(nb=nb, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    s.set_front_split(1 << 30)
    st = s.structure()
    n = st["supernode_size"] + st["work_size"]
    assert (n > 64).sum() > 0
    _check_fp64(s, p, name)
    t = s.front_times()
    assert np.all(t[n > 64] > 0) and np.all(t[n <= 64] == -1.0)
    xt, b = P.rhs(p, 1, seed=5)
    for prec in (mfdag.FP32_IR, mfdag.FP16_IR, mfdag.BF16_IR):
        assert s.factorize(p.val, prec) == 0
        x = s.solve(b.copy())
        eta = P.backward_error(p, x, b)
        log(f"{name} precision {prec}: eta {eta:.2e}, form {s.stats()['precision']}")
        assert eta <= 1e-14
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_medium_and_large_against_oracle(seed):
    """Per-node choices at random: children and parents of either kind."""
    p = P.weighted_laplacian_3d(18, seed=9)
    s = mfdag.Solver(nb=32, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 128, p.perm)
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 2, s.num_node).astype(np.int32)
    s.set_front_kinds(kinds)
    st = s.structure()
    n = st["supernode_size"] + st["work_size"]
    par = st["parent"] - 1
    big = n > 64
    mixed = [(k, par[k]) for k in range(s.num_node) if par[k] >= 0 and big[k] and big[par[k]] and kinds[k] != kinds[par[k]]]
    assert len(mixed) >= 2
    _check_fp64(s, p, f"mixed seed {seed}")
    s.close()


@pytest.mark.gpu
def test_medium_batch_replays_and_matches_oracle():
    systems = P.batch_anisotropic(40, seed=11, lo=16, hi=96)
    bp = P.batch_problem(systems)
    s = mfdag.Solver(nb=64, graph=True)
    s.analyze_batch([(q.ccs_col, q.ccs_row) for q in systems], 4096, [q.perm for q in systems])
    s.set_front_split(1 << 30)
    assert s.factorize(bp.val) == 0
    f, st, info = oracle_for(s, bp.val)
    assert info == 0
    errs = front_errors(s, f, st)
    assert errs.max() <= 1e-12, errs.max()
    ref = [s.front(k + 1).copy() for k in range(s.num_node)]
    for _ in range(3):  # graph replays, eager work in between
        s.solve(np.ones((bp.n, 1)))
        assert s.factorize(bp.val) == 0
        assert all(np.array_equal(s.front(k + 1), ref[k]) for k in range(s.num_node))
    e = mfdag.Solver(nb=64, graph=False)
    e.analyze_batch([(q.ccs_col, q.ccs_row) for q in systems], 4096, [q.perm for q in systems])
    e.set_front_split(1 << 30)
    assert e.factorize(bp.val) == 0
    assert all(np.array_equal(e.front(k + 1), ref[k]) for k in range(s.num_node))
    s.close()
    e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,max_zero,nb", CASES, ids=[c[0] for c in CASES])
def test_wide_schur_tiles_against_oracle(name, make, max_zero, nb):
    """Every Schur update (with its extend-add) in 128 x 128 tiles."""
    p = make()
    s = mfdag.Solver(nb=nb, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    s.tile_bound = 1
    assert s.tile_bound == 1
    _check_fp64(s, p, name + " wide")
    xt, b = P.rhs(p, 1, seed=6)
    for prec in (mfdag.FP32_IR, mfdag.FP16_IR):
        assert s.factorize(p.val, prec) == 0
        assert P.backward_error(p, s.solve(b.copy()), b) <= 1e-14
    s.close()
