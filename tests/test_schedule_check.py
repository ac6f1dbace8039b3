"""Static check of the factorization schedule (mfd_check_schedule): the
launches on their streams with their event edges as a DAG, and every task's
accesses to front entries, stored diagonal-block inverses and parent
entries; a conflict is a pair of tasks touching a common entry, one
writing, with no order between them (one launch, or two launches no DAG
path orders).  CPU only: the schedule is built on the host.

The schedules the library builds have no conflict, for both level rules,
every update period, front-kind choice, 128 x 128 Schur tiles and the
hybrid's host subtrees.  Two modeled schedules that it does not build show
what the check finds: the Schur tiles of fronts whose chain ends early
moved to a fourth stream forked after the launch with the front's last
diagonal task (variant 1: they read the rows that the front's last TRSM
tiles, one step launch later, still write), and forked after the front's
last step task (variant 2: their extend-add into the parent runs beside a
sibling's on the main stream, both adding into the same parent entries)."""
import re

import numpy as np
import pytest

import mfdag
from mfdag import problems as P


def solver(p, max_zero):
    s = mfdag.Solver(nb=64, host_only=True)
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    return s


CASES = [
    ("lap2d_100", lambda: P.laplacian_2d(100), 64),
    ("wlap3d_24", lambda: P.weighted_laplacian_3d(24, seed=1), 16384),
    ("st27_12", lambda: P.stencil27_3d(12, seed=3), 1024),
    ("blk3_8", lambda: P.block_stencil_3dof(8, seed=7), 256),
]


@pytest.mark.parametrize("name,make,max_zero", CASES, ids=[c[0] for c in CASES])
def test_built_schedules_have_no_conflict(name, make, max_zero):
    s = solver(make(), max_zero)
    for rule in (0, 1):  # levels from the leaves up, and from the roots down
        s.level_rule = rule
        for period in (1, 2, 3, 4, 0):
            s.update_period = period
            n, rep = s.check_schedule(0)
            assert n == 0, (rule, period, rep)
    s.update_period = 2
    s.tile_bound = 1  # every Schur update in 128 x 128 tiles over balanced K segments
    assert s.check_schedule(0)[0] == 0
    s.tile_bound = 0
    st = s.structure()
    sub, bins = mfdag.tree_pruning(st["supernode_size"], st["work_size"], st["parent"], 0.5, 4)
    s.set_host_subtrees(np.sort(sub[bins == 0] - 1))  # the hybrid: these fronts get no launch
    assert s.check_schedule(0)[0] == 0
    s.set_host_subtrees([])
    s.set_front_split(1 << 30)  # every front of order > 64 by one workgroup
    assert s.check_schedule(0)[0] == 0
    s.set_front_split(0)
    kinds = np.random.default_rng(5).integers(-1, 2, s.num_node).astype(np.int32)
    s.set_front_kinds(kinds)  # mixed per node
    assert s.check_schedule(0)[0] == 0
    s.close()


def test_headline_and_batch_schedules_have_no_conflict():
    s = solver(P.weighted_laplacian_3d(64, seed=2024), 16384)
    assert s.check_schedule(0) == (0, "")
    s.close()
    systems = P.batch_anisotropic(120, seed=3)
    b = mfdag.Solver(nb=64, host_only=True)
    b.analyze_batch([(q.ccs_col, q.ccs_row) for q in systems], 4096, [q.perm for q in systems])
    assert b.check_schedule(0)[0] == 0
    b.close()


def kinds_of(report):
    return {(int(a), wa, int(b), wb) for a, wa, b, wb in
            re.findall(r"kinds (-?\d+) \((\w)\) / (-?\d+) \((\w)\)", report)}


def test_modeled_fourth_stream_schedules_are_caught():
    s = solver(P.laplacian_2d(100), 64)
    n1, rep1 = s.check_schedule(1)
    assert n1 > 0
    # a step launch writes what a fourth-stream Schur tile reads
    assert all(k[1] == "w" and k[2] == -2 and k[3] == "r" and k[0] in (3, 4) for k in kinds_of(rep1)), rep1
    assert s.check_schedule(2)[0] == 0  # forked after the last step task: ordered here
    s.close()
    s = solver(P.weighted_laplacian_3d(24, seed=1), 16384)
    n2, rep2 = s.check_schedule(2)
    assert n2 > 0
    # a main-stream Schur tile and a fourth-stream one both add into one parent
    assert {frozenset((k[0], k[2])) for k in kinds_of(rep2)} == {frozenset((6, -2))}, rep2
    assert all(k[1] == "w" and k[3] == "w" for k in kinds_of(rep2))
    s.close()


def test_arguments():
    s = mfdag.Solver(nb=64, host_only=True)
    with pytest.raises(mfdag.MfdError) as e:
        s.check_schedule(0)
    assert e.value.code == mfdag.MFD_ERR_STATE
    p = P.laplacian_2d(12)
    s.analyze(p.ccs_col, p.ccs_row, 8, p.perm)
    with pytest.raises(mfdag.MfdError):
        s.check_schedule(3)
    s.close()
