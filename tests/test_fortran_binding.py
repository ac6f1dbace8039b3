"""The Fortran module hppllt (multifrontal-dag_amd/fortran/hppllt.f90) called
from a Fortran program through ISO_C_BINDING, as a caller of the reference
module would call it (tests/fortran/hppllt_check.f90).

CPU: the analysis getters the Fortran caller sees equal the C-ABI's on the
same matrix and orderings (natural, a caller ordering with use_metis = 1, and
use_metis = 1 without one: the library's nested dissection).
GPU: the same program factorizes and solves (FP64, 3 right-hand sides,
FP32 + refinement) and checks x against x_true."""
import os
import subprocess

import numpy as np
import pytest

import mfdag

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "multifrontal-dag_amd", "lib", "hppllt_check")


def run_check(k, nb, max_zero, mode, timeout=120):
    if not os.path.exists(CHECK):
        pytest.fail("hppllt_check is not built (run __graft_entry__.build())")
    r = subprocess.run([CHECK, str(k), str(nb), str(max_zero), mode], capture_output=True, text=True,
                       timeout=timeout)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if len(parts) == 2:
            out[parts[0]] = parts[1]
    assert r.stdout.strip().endswith("ok")
    return out


def laplacian_ccs(k):
    """The lower CCS the Fortran program builds (diagonal first)."""
    n = k * k
    col_ptr, rows = [1], []
    for j in range(k):
        for i in range(k):
            c = i + j * k + 1
            rows.append(c)
            if i < k - 1:
                rows.append(c + 1)
            if j < k - 1:
                rows.append(c + k)
            col_ptr.append(len(rows) + 1)
    return n, np.array(col_ptr, np.int32), np.array(rows, np.int32)


def red_black(k):
    idx = [c + 1 for r in (0, 1) for c in range(k * k) if ((c % k + 1) + (c // k + 1)) % 2 == r]
    return np.array(idx, np.int32)


@pytest.mark.parametrize("k,nb,max_zero", [(12, 4, 4), (20, 8, 0), (9, 3, 16)])
def test_fortran_analysis_matches_capi(k, nb, max_zero):
    out = run_check(k, nb, max_zero, "analyze")
    n, cp, rw = laplacian_ccs(k)
    assert int(out["n"]) == n and int(out["nnz"]) == len(rw)
    s = mfdag.Solver(nb=nb, host_only=True)
    s.analyze(cp, rw, max_zero, None)
    assert int(out["cost"]) == s.cost()
    assert int(out["nonzero"]) == s.nonzero()
    for w in (1, 2, 3):
        assert int(out[f"array_size_{w}"]) == s.array_size(w)
    s.close()
    s = mfdag.Solver(nb=nb, host_only=True)
    s.analyze(cp, rw, max_zero, red_black(k))
    assert int(out["perm_cost"]) == s.cost()
    s.close()
    # use_metis = 1 without a caller ordering: mfd_order_nd's
    s = mfdag.Solver(nb=nb, host_only=True)
    s.analyze(cp, rw, max_zero, mfdag.order_nd(cp, rw))
    assert int(out["nd_cost"]) == s.cost()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("k,nb,max_zero", [(30, 8, 8), (70, 64, 256)])
def test_fortran_factorize_solve(k, nb, max_zero):
    out = run_check(k, nb, max_zero, "solve")
    assert float(out["fp64_x_err"]) <= 1e-12
    assert float(out["multi_x_err"]) <= 1e-12
    assert 0 <= float(out["fp32ir_backward_error"]) <= 1e-14
    assert float(out["perm_x_err"]) <= 1e-12
    assert float(out["nd_x_err"]) <= 1e-12
