"""Pins the C oracle (oracle/mfd_oracle.c) against the reference's own known-answer
tests, transcribed in tests/golden/reference_kats.json, and against what the
mathematics fixes (LAPACK via numpy on random SPD tiles, dense Cholesky of
whole problems)."""
import numpy as np
import pytest

from oracle import oracle as O


def _arr(d, n):
    a = np.zeros(n)
    for k, v in d.items():
        a[int(k) - 1] = v
    return a


def test_tile_kernel_kats(kats):
    L = O.lib()
    t = kats["tile_kernels"]
    a = _arr(t["dpotrf"]["a"], 9)
    info = O.C.c_int()
    L.orc_dpotrf_u(3, a, O.C.byref(info))
    assert info.value == 0
    assert np.array_equal(a[np.array(t["dpotrf"]["pos"]) - 1], t["dpotrf"]["check"])

    d = _arr(t["dtrsm"]["diag"], 9)
    b = np.array(t["dtrsm"]["b"], float)
    L.orc_dtrsm_lutn(3, 2, d, b)
    assert np.array_equal(b, t["dtrsm"]["check"])

    a = np.array(t["dsyrk"]["a"], float)
    c = np.ones(4)
    L.orc_dsyrk_ut(2, 3, a, c)
    assert np.array_equal(c[np.array(t["dsyrk"]["pos"]) - 1], t["dsyrk"]["check"])

    g = t["dgemm"]
    c = np.ones(8)
    L.orc_dgemm_tn(g["upper_n"], g["lower_n"], g["ld"], np.array(g["upper"], float), np.array(g["lower"], float), c)
    assert np.array_equal(c, g["check"])

    u = np.zeros(9)
    for i in range(1, 4):
        for j in range(i, 4):
            u[(j - 1) * 3 + i - 1] = 1.0
    x = np.array(t["dtrsv_l"]["b"], float)
    L.orc_dtrsv_ut(3, u, x)
    assert np.array_equal(x, t["dtrsv_l"]["check"])
    x = np.array(t["dtrsv_u"]["b"], float)
    L.orc_dtrsv_un(3, u, x)
    assert np.array_equal(x, t["dtrsv_u"]["check"])

    y = np.ones(2)
    L.orc_dgemv_t(3, 2, np.arange(1, 7, dtype=float), np.array([1.0, 2, 3]), y)
    assert np.array_equal(y, t["dgemv_t"]["check"])
    y = np.ones(3)
    L.orc_dgemv_n(3, 2, np.arange(1, 7, dtype=float), np.array([1.0, 2]), y)
    assert np.array_equal(y, t["dgemv_n"]["check"])


def test_border_kernel_kats(kats):
    """border_potrf / border_trsm compose exactly as border_kernel_m.f90:9-30."""
    L = O.lib()
    t = kats["tile_kernels"]["border_potrf"]
    a = np.array(t["a"], float)
    b = np.array(t["b"], float)
    ss, ws = t["ssize"], t["wsize"]
    info = O.C.c_int()
    head = a[: ss * ss].copy()
    L.orc_dpotrf_u(ss, head, O.C.byref(info))
    a[: ss * ss] = head
    rect = a[ss * ss:].copy()
    L.orc_dtrsm_lutn(ss, ws, head, rect)
    L.orc_dsyrk_ut(ws, ss, rect, b)
    a[ss * ss:] = rect
    assert np.array_equal(a[np.array(t["a_pos"]) - 1], t["a_check"])
    assert np.array_equal(b[np.array(t["b_pos"]) - 1], t["b_check"])

    for key in ("border_trsm_1", "border_trsm_2"):
        t = kats["tile_kernels"][key]
        ss, ws, nrow = t["ssize"], t["wsize"], t["nrow"]
        diag = np.array(t["diag"], float)
        lo = np.array(t["lower_a"], float)
        lb = np.zeros(ws * nrow)
        L.orc_dtrsm_lutn(ss, nrow, diag[: ss * ss].copy(), lo)
        L.orc_dgemm_tn(ws, nrow, ss, diag[ss * ss:].copy(), lo, lb)
        assert np.array_equal(lo, t["check_a"]), key
        assert np.array_equal(lb, t["check_b"]), key


def test_extend_add_kats(kats):
    L = O.lib()
    t = kats["tile_kernels"]["extend_add_tri"]
    to = np.full(81, float(t["to_init"]))
    loc = np.array(t["local"], np.int32)
    L.orc_extend_add_tri(np.array(t["from"], float), to, loc, len(loc), t["ld_from"], t["ld_to"], t["offset"])
    for r, q, v in t["check"]:
        assert to[r - 1 + (q - 1) * 9] == v
    for key in ("extend_add_rect1", "extend_add_rect2"):
        t = kats["tile_kernels"][key]
        to = np.full(81, float(t["to_init"]))
        cl = np.array(t["col_local"], np.int32)
        rl = np.array(t["row_local"], np.int32)
        L.orc_extend_add_rect(np.array(t["from"], float), to, cl, len(cl), rl, len(rl), t["ld_from"], t["ld_to"],
                              t["coffset"], t["roffset"])
        for r, q, v in t["check"]:
            assert to[r - 1 + (q - 1) * 9] == v, (key, r, q)


def _three_front(kats, nb):
    t = kats["three_front"]
    f = O.OracleFactor(t["supernode_size"], t["work_size"], t["parent"], nb, t["local_index"])
    lens = np.array(t["supernodal_index_lengths"], np.int64)
    colptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    f.assemble(colptr, t["supernodal_index"], np.array(t["val"], float))
    return f


def check_three_front_factors(kats, nb, get_tile):
    """Asserts the reference's expected integer tiles (seq_factorize_test.f90)
    on whatever `get_tile(node, i, j)` returns (flat supernode tile)."""
    for chk in kats["three_front"]["factor_checks"][str(nb)]:
        kind, node, i, j = chk[:4]
        tile = np.asarray(get_tile(node, i, j))
        if kind == "full":
            assert np.array_equal(tile, chk[4]), chk
        elif kind in ("tri", "tri_head"):
            n = chk[5]
            sub = tile[: n * n].reshape(n, n, order="F")
            vals = [sub[jj, ii] for ii in range(n) for jj in range(ii + 1)]
            assert np.array_equal(vals, chk[4]), chk
        elif kind == "slice":
            lo, hi = chk[5], chk[6]
            assert np.array_equal(tile[lo - 1:hi], chk[4]), chk


@pytest.mark.parametrize("nb", [2, 3])
def test_three_front_factorize(kats, nb):
    f = _three_front(kats, nb)
    assert f.factorize() == 0
    check_three_front_factors(kats, nb, lambda s, i, j: f.tile_view(1, s, i, j))


@pytest.mark.parametrize("nb", [2, 3])
def test_three_front_forward_backward(kats, nb):
    t = kats["three_front"]
    for which, rhs in (("forward", t["forward_rhs"]), ("backward", t["backward_rhs"])):
        f = O.OracleFactor(t["supernode_size"], t["work_size"], t["parent"], nb, t["local_index"])
        for node, i, j, vals in t["solve_tiles"][str(nb)]:
            # the Fortran test assigns a 4-element literal to the 3-element
            # tile (3,2,1) at nb = 3; the extra element lands in the next tile,
            # which the following assignment overwrites, as here
            off, _, _ = f.tile(1, node, i, j)
            f.array(1)[off:off + len(vals)] = vals
        out = getattr(f, which)(np.array(rhs, float))
        assert np.array_equal(out, t["solve_expected"]), (which, nb)


def test_block_local_index_kats(kats):
    for case in kats["block_local_index_info"]["cases"]:
        parent = [0] * len(case["supernode_size"])
        f = O.OracleFactor(case["supernode_size"], case["work_size"], parent, case["nb"], case["local_index"])
        for node, exp in case["groups"].items():
            got = f.groups(int(node))
            assert len(got) == len(exp), (node, got, exp)
            for g, e in zip(got, exp):
                for a, b in zip(g, e):
                    if b is not None:
                        assert a == b, (node, g, e)


def test_cost_kats(kats):
    from oracle import analysis as A
    for case in kats["cost"]["cases"]:
        assert A.cost(case["supernode_size"], case["work_size"]) == case["cost"]
        if "nonzero" in case:
            assert A.nonzero(case["supernode_size"], case["work_size"]) == case["nonzero"]


def test_tile_kernels_match_lapack():
    """Random SPD tiles: the oracle's textbook kernels against numpy's LAPACK."""
    L = O.lib()
    rng = np.random.default_rng(1)
    for n in (1, 5, 17, 40):
        m = rng.standard_normal((n, n))
        spd = m @ m.T + n * np.eye(n)
        a = spd.flatten(order="F").copy()
        info = O.C.c_int()
        L.orc_dpotrf_u(n, a, O.C.byref(info))
        u = np.triu(a.reshape(n, n, order="F"))
        assert np.allclose(u, np.linalg.cholesky(spd).T, rtol=1e-12, atol=1e-12)
        k = 7
        b = rng.standard_normal((n, k))
        x = b.flatten(order="F").copy()
        L.orc_dtrsm_lutn(n, k, u.flatten(order="F").copy(), x)
        assert np.allclose(u.T @ x.reshape(n, k, order="F"), b, rtol=1e-11, atol=1e-11)
    # a non-SPD tile reports its failing column
    a = np.array([1.0, 2.0, 2.0, 1.0])
    info = O.C.c_int()
    L.orc_dpotrf_u(2, a, O.C.byref(info))
    assert info.value == 2
