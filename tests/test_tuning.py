"""Schedule tuner: the in-panel update period (mfd_set_update_period), the
cost model's HIP kernels (mfd_mlp_*) and the tuner built on them
(mfdag/tune.py).  None of it has a counterpart in the reference.

CPU: the numpy restatement of the model (oracle/mlp.py) is pinned by what
the mathematics fixes -- central differences of the loss, the closed forms
of a network with zero output weights, Adam's first step -- and the C-ABI's
argument checks (no compute without a GPU: valid calls report no device).
GPU: the kernels against the restatement (prediction, gradient, 300 Adam
steps), every period giving the oracle's factor, and the tuner: records
from the exhaustive sweep, the model trained on some problems and measured by
its regret on others."""
import sys

import numpy as np
import pytest

import mfdag
from mfdag import problems as P
from mfdag import tune
from oracle import mlp as M

from parity import front_errors, oracle_for


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sample(seed, n=300):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.5, M.N_PAR)
    x = rng.normal(0.0, 1.0, (n, M.N_IN))
    t = np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.2 * x[:, 3]
    return p, x, t


# ------------------------------------------------------------------- CPU
def test_oracle_gradient_against_central_differences():
    p, x, t = sample(1, 40)
    g, loss = M.gradient(p, x, t)
    assert loss == pytest.approx(M.loss(p, x, t), rel=1e-14)
    h = 1e-6
    fd = np.empty(M.N_PAR)
    for k in range(M.N_PAR):
        e = np.zeros(M.N_PAR)
        e[k] = h
        fd[k] = (M.loss(p + e, x, t) - M.loss(p - e, x, t)) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(g).max())


def test_oracle_closed_forms():
    p, x, t = sample(2, 50)
    W1, b1, w2, b2 = M.unpack(p)
    w2[:] = 0.0
    y = M.predict(p, x)
    assert np.all(y == b2)
    g, loss = M.gradient(p, x, t)
    assert loss == pytest.approx(np.mean((b2 - t) ** 2), rel=1e-13)
    assert g[-1] == pytest.approx(2.0 * (b2 - t.mean()), rel=1e-12)
    assert np.all(g[:M.N_HID * M.N_IN + M.N_HID] == 0.0)
    # the output weights' gradient is 2 mean((b2 - t) h_j)
    h = np.tanh(x @ W1.T + b1)
    assert np.allclose(g[M.N_HID * M.N_IN + M.N_HID:-1], 2.0 * ((b2 - t) @ h) / len(x), rtol=1e-12, atol=0)


def test_oracle_adam_first_step():
    p, x, t = sample(3, 30)
    g, _ = M.gradient(p, x, t)
    q, _ = M.train(p, x, t, 1, 1e-3)
    # m_hat = g, v_hat = g^2: each parameter moves by lr g / (|g| + eps)
    assert np.allclose(p - q, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-18)


def test_arguments_checked_without_compute():
    L = mfdag.lib()
    p, x, t = sample(4, 10)
    assert mfdag.MLP_PARAMS == M.N_PAR
    with pytest.raises(mfdag.MfdError) as e:
        mfdag.mlp_train(p, x[:0], t[:0], 10, 0.01)
    assert e.value.code == mfdag.MFD_ERR_ARG
    with pytest.raises(mfdag.MfdError) as e:
        mfdag.mlp_train(p, x, t, 10, 0.0)
    assert e.value.code == mfdag.MFD_ERR_ARG
    s = mfdag.Solver(host_only=True)
    assert s.update_period == 2
    for bad in (-1, 9):
        with pytest.raises(mfdag.MfdError):
            s.update_period = bad
    s.update_period = 3
    assert s.update_period == 3
    q = mfdag.problems.laplacian_2d(12)
    s.analyze(q.ccs_col, q.ccs_row, 4, q.perm)
    s.update_period = 1
    assert s.update_period == 1
    s.close()
    assert L.mfd_get_update_period(None) == mfdag.MFD_ERR_ARG
    if not s.device:
        # no GPU here: a valid call must report that, not compute on the host
        with pytest.raises(mfdag.MfdError) as e:
            mfdag.mlp_predict(p, x)
        assert e.value.code == mfdag.MFD_ERR_NODEVICE


def test_regret_and_gain_bookkeeping():
    class Fixed:
        def __init__(self, p):
            self.p = p

        def choose(self, n, w, elem_bytes, periods):
            assert periods == (1, 2, 3)
            return self.p

    recs = [dict(problem="a", precision=0, period=k, n=900, w=0, elem_bytes=8, ms=ms)
            for k, ms in ((1, 10.0), (2, 8.0), (3, 9.0))]
    recs += [dict(problem="a", precision=1, period=k, n=900, w=0, elem_bytes=4, ms=ms)
             for k, ms in ((1, 5.0), (2, 6.0), (3, 7.0))]
    r = tune.regret(Fixed(3), recs)
    assert [(q["precision"], q["best_period"]) for q in r] == [(0, 2), (1, 1)]
    assert r[0]["regret"] == pytest.approx(9.0 / 8.0 - 1.0) and r[0]["gain"] == pytest.approx(8.0 / 9.0)
    assert r[1]["regret"] == pytest.approx(0.4) and r[1]["gain"] == pytest.approx(6.0 / 7.0)
    assert tune.features(1024, 3, 0, 8) == [10.0, 3.0, 0.0, 2.0]
    assert tune.pick({1: 10.0, 2: 10.04, 3: 11.0}) == 2      # within 0.5%: the default stays
    assert tune.pick({1: 9.9, 2: 10.0, 3: 11.0}) == 1        # 1% faster: moves
    assert tune.pick({1: 9.9, 3: 11.0}) == 1


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_mlp_kernels_against_oracle():
    p, x, t = sample(5, 1000)  # eight gradient workgroups, the last one partial
    y = mfdag.mlp_predict(p, x)
    assert np.abs(y - M.predict(p, x)).max() <= 1e-13 * np.abs(y).max()
    g, loss = mfdag.mlp_gradient(p, x, t)
    go, lo = M.gradient(p, x, t)
    log(f"gradient rel err {np.abs(g - go).max() / np.abs(go).max():.2e}")
    assert np.abs(g - go).max() <= 1e-13 * np.abs(go).max()
    assert loss == pytest.approx(lo, rel=1e-13)
    q, lq = mfdag.mlp_train(p, x, t, 300, 0.01)
    qo, lqo = M.train(p, x, t, 300, 0.01)
    log(f"trained params rel err {np.abs(q - qo).max() / np.abs(qo).max():.2e}, loss {lo:.3e} -> {lq:.3e}")
    assert np.abs(q - qo).max() <= 1e-9 * np.abs(qo).max()
    assert lq == pytest.approx(lqo, rel=1e-9)
    assert lq < 0.5 * lo
    # deterministic: the same call twice gives the same bits
    q2, _ = mfdag.mlp_train(p, x, t, 300, 0.01)
    assert np.array_equal(q, q2)


@pytest.mark.gpu
def test_every_update_period_gives_the_oracle_factor():
    p = P.weighted_laplacian_3d(20, seed=3)
    s = mfdag.Solver(nb=64, graph=True)
    s.analyze(p.ccs_col, p.ccs_row, 16384, p.perm)
    n, c, w = tune.largest_front(s)
    assert c >= 5 * 64  # enough 64-column steps for periods 1..4
    f, st, info = oracle_for(s, p.val)
    assert info == 0
    xt, b = P.rhs(p, 2, seed=4)
    for period in (1, 3, 4, 8, 0, 2):
        s.update_period = period
        with pytest.raises(mfdag.MfdError):  # the schedule changed: factor again first
            s.solve(b)
        for prec in (mfdag.FP64, mfdag.FP32_IR):
            assert s.factorize(p.val, prec) == 0
            x = s.solve(b)
            eta = P.backward_error(p, x, b)
            log(f"period {period} precision {prec}: eta {eta:.2e}")
            assert eta <= 1e-14
        assert s.factorize(p.val, mfdag.FP64) == 0
        errs = front_errors(s, f, st)
        assert errs.max() <= 1e-12, (period, errs.max())
    s.close()


TRAIN = [("wlap3d_%d" % k, lambda k=k: P.weighted_laplacian_3d(k, seed=k)) for k in (16, 24, 32, 40, 48, 56, 60)] + \
        [("st27_%d" % k, lambda k=k: P.stencil27_3d(k, seed=k)) for k in (12, 20, 28, 36, 44)] + \
        [("lap2d_%d" % k, lambda k=k: P.laplacian_2d(k)) for k in (100, 200)]
HELD_OUT = [("wlap3d_36", lambda: P.weighted_laplacian_3d(36, seed=36)),
            ("wlap3d_52", lambda: P.weighted_laplacian_3d(52, seed=52)),
            ("wlap3d_64", lambda: P.weighted_laplacian_3d(64, seed=2024)),
            ("st27_24", lambda: P.stencil27_3d(24, seed=24)),
            ("st27_40", lambda: P.stencil27_3d(40, seed=40)),
            ("lap2d_160", lambda: P.laplacian_2d(160))]


def sweep(cases):
    recs = []
    for name, make in cases:
        p = make()
        s = mfdag.Solver(nb=64, graph=True)
        s.analyze(p.ccs_col, p.ccs_row, 16384, p.perm)
        xt, b = P.rhs(p, 1, seed=1)
        r = tune.record(name, s, p.val, b[:, 0], precisions=tune.MODEL_FORMS)
        log(name, [(q["precision"], q["period"], round(q["ms"], 3)) for q in r])
        recs += r
        s.close()
    return recs


@pytest.mark.gpu
def test_tuner_cost_model_regret():
    train = sweep(TRAIN)
    held = sweep(HELD_OUT)
    assert all(r["ms"] > 0 for r in train + held)
    model = tune.CostModel(seed=0)
    x, t0 = model.scale(train)
    loss0 = float(np.mean((mfdag.mlp_predict(model.params, x) - t0) ** 2))
    model.fit(train)
    log(f"cost model loss {loss0:.3f} -> {model.loss:.4f} on {len(train)} records")
    assert model.loss < 0.1 * loss0
    res = tune.regret(model, held)
    for r in res:
        log(f"{r['problem']} prec {r['precision']}: chose {r['chosen']} (best {r['best_period']}), "
            f"regret {100 * r['regret']:.1f}%, gain over period {tune.DEFAULT_PERIOD} {r['gain']:.3f}x")
    mean_regret = float(np.mean([r["regret"] for r in res]))
    mean_gain = float(np.exp(np.mean(np.log([r["gain"] for r in res]))))
    log(f"held-out: mean regret {100 * mean_regret:.1f}%, geometric-mean gain {mean_gain:.3f}x")

    class Worst:
        """Always the slowest period of the group: what a period-blind
        choice can be unlucky enough to get."""
        def __init__(self, recs):
            self.t = {}
            for r in recs:
                self.t.setdefault((r["n"], r["w"], r["elem_bytes"]), {})[r["period"]] = r["ms"]

        def choose(self, n, w, elem_bytes, periods):
            t = self.t[(n, w, elem_bytes)]
            return max(periods, key=lambda p: t[p])

    worst_regret = float(np.mean([r["regret"] for r in tune.regret(Worst(held), held)]))
    # control: the same model trained with the period feature shuffled cannot
    # tell the periods apart
    rng = np.random.default_rng(1)
    shuffled = [dict(r) for r in train]
    per = rng.permutation([r["x"][1] for r in train])
    for r, v in zip(shuffled, per):
        r["x"] = [r["x"][0], float(v), r["x"][2], r["x"][3]]
    blind = tune.CostModel(seed=0).fit(shuffled)
    blind_regret = float(np.mean([r["regret"] for r in tune.regret(blind, held)]))
    log(f"held-out: always-slowest mean regret {100 * worst_regret:.1f}%, "
        f"period-shuffled model {100 * blind_regret:.1f}%")
    assert mean_regret <= 0.01
    assert mean_gain >= 1.0
    assert mean_regret <= 0.5 * worst_regret
    assert mean_regret <= blind_regret


# ----------------------------------------------- per-front medium/large choice
def _open(make, max_zero, profile=False):
    p = make()
    s = mfdag.Solver(nb=64, graph=not profile, profile=profile)
    if isinstance(p, list):  # a batch of systems
        bp = P.batch_problem(p)
        s.analyze_batch([(q.ccs_col, q.ccs_row) for q in p], max_zero, [q.perm for q in p])
        return s, bp.val, p
    s.analyze(p.ccs_col, p.ccs_row, max_zero, p.perm)
    return s, p.val, p


FRONT_TRAIN = [("wlap3d_24", lambda: P.weighted_laplacian_3d(24, seed=24), 16384),
               ("lap2d_200", lambda: P.laplacian_2d(200), 4096),
               ("st27_16", lambda: P.stencil27_3d(16, seed=16), 4096),
               ("blk3_12", lambda: P.block_stencil_3dof(12, seed=12), 4096),
               ("batch_96", lambda: P.batch_anisotropic(96, seed=5), 4096)]
FRONT_HELD = [("lap2d_400", lambda: P.laplacian_2d(400), 4096),
              ("wlap3d_40", lambda: P.weighted_laplacian_3d(40, seed=40), 16384),
              ("batch_192", lambda: P.batch_anisotropic(192, seed=9), 4096)]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [mfdag.FP64, mfdag.FP32_IR])
def test_front_policy_regret(prec):
    """The per-front medium/large choice (tune.choose_kinds: the perceptron
    trained on the medium kernel's per-front clocks, the step and Schur
    launches by constants fitted to profiled runs) against the exhaustive
    tuner over global pivot bounds, on problems outside the training set.
    Controls: every front medium (the blind extreme) does far worse, and the
    perceptron fitted to targets shuffled across fronts does not fit."""
    eb = 8 if prec == mfdag.FP64 else 4
    recs, crecs = [], []
    for name, make, mz in FRONT_TRAIN:
        s, a, p = _open(make, mz)
        recs.append(tune.front_records(s, a, prec))
        s.close()
        if isinstance(p, list):
            crecs.append(tune.chain_record(None, None, mz, [q.perm for q in p], a, prec,
                                           batch=[(q.ccs_col, q.ccs_row) for q in p]))
        else:
            crecs.append(tune.chain_record(p.ccs_col, p.ccs_row, mz, p.perm, a, prec))
    tau, t0, rate = tune.chain_constants(crecs)
    model = tune.FrontModel(seed=0).fit(recs)
    max_order = int(max(r["n"].max() for r in recs))
    log(f"prec {prec}: {sum(len(r['us']) for r in recs)} front records, loss {model.loss:.4f}; "
        f"step round {tau:.1f} us, Schur {t0:.1f} us/launch + {rate / 1e6:.1f} TFLOP/s")
    shuffled = [dict(r, us=np.random.default_rng(3).permutation(r["us"])) for r in recs]
    blind = tune.FrontModel(seed=0).fit(shuffled)
    assert blind.loss > 3 * model.loss
    regrets, extreme = [], []
    for name, make, mz in FRONT_HELD:
        s, a, p = _open(make, mz)
        t = tune.measure_split(s, a, prec)
        kinds = tune.choose_kinds(s.structure(), model, tau, t0, rate, eb, max_order)
        tk = tune.measure_kinds(s, a, prec, kinds)
        best = min(t.values())
        regrets.append(tk / best - 1.0)
        extreme.append(t[max(t)] / best - 1.0)
        log(f"{name}: bounds {({k: round(v, 2) for k, v in t.items()})}, policy {tk:.2f} ms "
            f"({int(kinds.sum())} of {s.num_node} fronts medium), regret {100 * regrets[-1]:.1f}%")
        s.close()
    log(f"mean regret {100 * np.mean(regrets):.1f}%, every front medium {100 * np.mean(extreme):.0f}%")
    assert np.mean(regrets) <= 0.03
    assert max(regrets) <= 0.10
    assert np.mean(extreme) >= 10 * max(np.mean(regrets), 0.01)
