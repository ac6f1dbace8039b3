"""MatrixMarket input and the command-line driver, on the reference's own
end-to-end matrices (hppllt_test.f90:12-15: test_matrix/bcsstk01.mtx and
test_matrix/poisson_9_9.mtx, options nb = 10, max_zero = 10, copied under
tests/golden/ as data).

CPU: the reader's CCS equals an independent COO -> CSC construction (scipy)
entry for entry; the column order of an entry given twice; error paths; the
analysis of both matrices equals the oracle's (use_metis = 0, the reference
test's second pair); the driver's record has example.f90:47-60's fields and
widths and the analysis getters' values.
GPU: the driver factorizes and solves both matrices (FP64: the reference
test's bound, error <= 1e-8; FP32 + refinement: the bound its backward-error
target implies); the same problems through the Python boundary give fronts
and x that agree with the oracle's."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import mfdag
from oracle import analysis as A
from parity import front_errors, oracle_for
from test_capi_host import compare_with_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "multifrontal-dag_amd", "lib", "mfdag_example")
MATRICES = ["bcsstk01.mtx", "poisson_9_9.mtx"]
# example.f90:47-60: i8 i8 i8 i12 i20 3*i20 7*f10.3 e10.3
WIDTHS = [8, 8, 8, 12, 20, 20, 20, 20] + [10] * 7 + [10]


def independent_ccs(path):
    """The file's lower triangle through scipy: CSC with sorted rows."""
    lines = [l for l in open(path) if not l.startswith("%")]
    n, m, nnz = map(int, lines[0].split())
    t = np.array([l.split() for l in lines[1:1 + nnz]], dtype=float)
    r, c = t[:, 0].astype(int) - 1, t[:, 1].astype(int) - 1
    M = sp.csc_matrix((t[:, 2], (r, c)), shape=(n, n))
    M.sort_indices()
    return n, M.indptr + 1, M.indices + 1, M.data, M


@pytest.mark.parametrize("name", MATRICES)
def test_reader_matches_scipy(name):
    path = os.path.join(GOLDEN, name)
    n, col, row, val = mfdag.read_matrix_market(path)
    n2, col2, row2, val2, _ = independent_ccs(path)
    assert n == n2
    assert np.array_equal(col, col2)
    assert np.array_equal(row, row2)
    assert np.array_equal(val, val2)
    # diagonal first in every column
    assert np.array_equal(row[col[:-1] - 1], np.arange(1, n + 1))


def write(tmp_path, text, name="m.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reader_order_and_forms(tmp_path):
    # no banner, comments before the size line, a D exponent, an entry given
    # twice (the later one comes first, as example.f90:104-143 leaves it),
    # columns given out of order
    text = "% comment\n%another\n3 3 6\n3 2 5.0\n1 1 1.5D+00\n2 1 -1\n2 2 4\n3 3 9e0\n3 2 7\n"
    n, col, row, val = mfdag.read_matrix_market(write(tmp_path, text))
    assert n == 3
    assert col.tolist() == [1, 3, 6, 7]
    assert row.tolist() == [1, 2, 2, 3, 3, 3]
    assert val.tolist() == [1.5, -1.0, 4.0, 7.0, 5.0, 9.0]
    # a banner in capitals with an integer field
    text = "%%MatrixMarket MATRIX Coordinate Integer Symmetric\n2 2 3\n1 1 2\n2 1 1\n2 2 2\n"
    n, col, row, val = mfdag.read_matrix_market(write(tmp_path, text))
    assert col.tolist() == [1, 3, 4] and row.tolist() == [1, 2, 2] and val.tolist() == [2, 1, 2]


@pytest.mark.parametrize("text,code", [
    ("2 2 2\n1 1 1\n1 2 1\n", mfdag.MFD_ERR_ARG),        # above the diagonal
    ("2 2 2\n1 1 1\n3 1 1\n", mfdag.MFD_ERR_ARG),        # row out of range
    ("2 2 2\n1 1 1\n1 0 1\n", mfdag.MFD_ERR_ARG),        # column 0
    ("2 2 3\n1 1 1\n2 2 1\n", mfdag.MFD_ERR_ARG),        # short file
    ("2 3 2\n1 1 1\n2 2 1\n", mfdag.MFD_ERR_ARG),        # not square
    ("2 2 2\n1 1 x\n2 2 1\n", mfdag.MFD_ERR_ARG),        # unreadable value
    ("%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", mfdag.MFD_ERR_UNSUPPORTED),
    ("%%MatrixMarket matrix coordinate pattern symmetric\n1 1 1\n1 1\n", mfdag.MFD_ERR_UNSUPPORTED),
    ("%%MatrixMarket matrix array real symmetric\n1 1\n1\n", mfdag.MFD_ERR_UNSUPPORTED),
])
def test_reader_errors(tmp_path, text, code):
    with pytest.raises(mfdag.MfdError) as e:
        mfdag.read_matrix_market(write(tmp_path, text))
    assert e.value.code == code


def test_reader_missing_file(tmp_path):
    with pytest.raises(mfdag.MfdError):
        mfdag.read_matrix_market(str(tmp_path / "absent.mtx"))


@pytest.mark.parametrize("name", MATRICES)
@pytest.mark.parametrize("max_zero", [10, 0, 3])
def test_analysis_matches_oracle(name, max_zero):
    n, col, row, _ = mfdag.read_matrix_market(os.path.join(GOLDEN, name))
    s, _ = compare_with_oracle(col, row, max_zero)
    s.close()


def run_cli(args, timeout=120):
    if not os.path.exists(CLI):
        pytest.fail("mfdag_example is not built (run __graft_entry__.build())")
    r = subprocess.run([CLI] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr)
    return r


def parse_record(path):
    line = open(path).read().rstrip("\n")
    assert len(line) == sum(WIDTHS), line
    fields, k = [], 0
    for w in WIDTHS:
        fields.append(line[k:k + w])
        k += w
    return fields


def check_record_structure(fields, name, max_zero, nb, natural):
    n, col, row, _ = mfdag.read_matrix_market(os.path.join(GOLDEN, name))
    s = mfdag.Solver(nb=nb, host_only=True)
    s.analyze(col, row, max_zero, None if natural else mfdag.order_nd(col, row))
    assert int(fields[0]) == max_zero and int(fields[1]) == nb and int(fields[2]) == 1
    assert int(fields[3]) == s.nonzero()
    assert int(fields[4]) == s.cost()
    for w in (1, 2, 3):
        assert int(fields[4 + w]) == s.array_size(w)
    s.close()


@pytest.mark.parametrize("name", MATRICES)
@pytest.mark.parametrize("natural", [True, False])
def test_cli_analyze_record(tmp_path, name, natural):
    rec = tmp_path / "rec.txt"
    r = run_cli([os.path.join(GOLDEN, name), 10, 10, rec, "--analyze-only"] + (["--natural"] if natural else []))
    assert r.returncode == 0
    assert r.stdout.split() == ["I", "A"]
    fields = parse_record(rec)
    check_record_structure(fields, name, 10, 10, natural)
    for f in fields[8:15]:
        assert float(f) >= 0.0
    assert fields[15].strip() == "NaN"


def test_cli_argument_errors(tmp_path):
    r = run_cli([os.path.join(GOLDEN, MATRICES[0]), 10, 10])
    assert r.returncode != 0 and "command line argument error" in r.stdout
    bad = write(tmp_path, "2 2 2\n1 1 1\n1 2 1\n")
    r = run_cli([bad, 10, 10, tmp_path / "rec.txt", "--analyze-only"])
    assert r.returncode != 0 and "code" in r.stderr
    perm = tmp_path / "perm.txt"
    perm.write_text("1 2 3\n")
    r = run_cli([os.path.join(GOLDEN, MATRICES[0]), 10, 10, tmp_path / "rec.txt", "--analyze-only", "--perm", perm])
    assert r.returncode != 0
    r = run_cli([os.path.join(GOLDEN, MATRICES[0]), 10, 10, tmp_path / "rec.txt", "--natural", "--perm", perm])
    assert r.returncode != 0 and "command line argument error" in r.stdout


def test_cli_caller_ordering(tmp_path):
    """--perm: the record's structure is that of the analysis with the
    ordering (use_metis = 1 with the caller's permutation)."""
    name = "poisson_9_9.mtx"
    n, col, row, _ = mfdag.read_matrix_market(os.path.join(GOLDEN, name))
    # red-black order of the 9 x 9 grid
    perm = [c + 1 for r in (0, 1) for c in range(n) if (c % 9 + c // 9) % 2 == r]
    pf = tmp_path / "perm.txt"
    pf.write_text("\n".join(map(str, perm)) + "\n")
    rec = tmp_path / "rec.txt"
    assert run_cli([os.path.join(GOLDEN, name), 4, 10, rec, "--analyze-only", "--perm", pf]).returncode == 0
    fields = parse_record(rec)
    exp = A.hppllt_analyze([list(row[col[j] - 1:col[j + 1] - 1]) for j in range(n)], 4, perm)
    assert int(fields[4]) == A.cost(exp["node_sizes"], exp["work_sizes"])
    assert int(fields[3]) == A.nonzero(exp["node_sizes"], exp["work_sizes"])


def example_error(M, x, rh):
    """example.f90:147-180 on the host: ||rh - A x||_2 / ||x||_2."""
    A_full = M + sp.tril(M, -1).T
    return np.linalg.norm(rh - A_full @ x) / np.linalg.norm(x)


@pytest.mark.gpu
@pytest.mark.parametrize("name", MATRICES)
@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("natural", [True, False])
def test_cli_end_to_end(tmp_path, name, fp32, natural):
    """hppllt_test.f90:12-15's four runs: options(4) = 0 (--natural) and 1
    (the library's ordering in place of METIS)."""
    rec = tmp_path / "rec.txt"
    args = [os.path.join(GOLDEN, name), 10, 10, rec] + (["--fp32"] if fp32 else []) + (["--natural"] if natural else [])
    r = run_cli(args)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["I", "A", "F", "S"]
    fields = parse_record(rec)
    check_record_structure(fields, name, 10, 10, natural)
    err = float(fields[15])
    if not fp32:
        # the reference test's bound (hppllt_test.f90:37)
        assert 0.0 <= err <= 1e-8, err
    else:
        # FP32 + refinement stops at a normwise backward error of 1e-14,
        # ||b - A x||_inf <= 1e-14 (||A||_inf ||x||_inf + ||b||_inf); with
        # ||b||_inf <= ||A||_inf ||x||_inf that bounds this unscaled measure
        # by 2e-14 sqrt(n) ||A||_inf (bcsstk01: ||A||_inf = 3.6e9).
        _, _, _, _, M = independent_ccs(os.path.join(GOLDEN, name))
        A_full = M + sp.tril(M, -1).T
        norm_a = abs(A_full).sum(axis=1).max()
        assert 0.0 <= err <= 2e-14 * np.sqrt(M.shape[0]) * norm_a, err


@pytest.mark.gpu
@pytest.mark.parametrize("name", MATRICES)
def test_reference_matrices_parity(name):
    """hppllt_test.f90:24-37 through the Python boundary: b = 10 in every
    entry; fronts and x against the oracle on the same structure."""
    path = os.path.join(GOLDEN, name)
    n, col, row, val = mfdag.read_matrix_market(path)
    _, _, _, _, M = independent_ccs(path)
    s = mfdag.Solver(nb=10)
    s.analyze(col, row, 10, None)
    s.factorize(val)
    f, st, info = oracle_for(s, val)
    assert info == 0
    assert front_errors(s, f, st).max() <= 1e-12
    b = np.full(n, 10.0)
    x = s.solve(b)
    perm = st["perm"] - 1
    xo = np.empty(n)
    xo[perm] = f.solve(b[perm])
    assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo).max()
    assert example_error(M, x, 10.0) <= 1e-8
    s.close()
