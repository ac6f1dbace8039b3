"""Schedule tuner for the large-front factorization.

The choice tuned is the in-panel update period of the large fronts
(mfd_set_update_period): every `period` 64-column steps the trailing
columns receive the last `period` panels as one update with K = 64 period,
and the steps in between carry the pending panels in their TRSM tiles and
diagonal tasks.  Short periods keep the critical-path tasks light; long ones
make fewer, deeper passes over the trailing columns.

Two ways to choose it:
  * measure (the exhaustive tuner): factor (and, given right-hand sides,
    solve) the analysed problem with every period and keep the fastest; it
    records one sample per (problem, precision, period);
  * CostModel: a perceptron (the mfd_mlp_* HIP kernels) trained on such
    records predicts the factor + solve time from the features of the
    problem's front with the most pivot columns -- order n, update width
    K (64 period), update rows w, bytes per entry -- and picks the period with the smallest
    prediction, without a sweep.  The measured tuner is its fallback and the
    yardstick of its regret (regret below).  The model covers the FP64 form
    (MODEL_FORMS): in the FP32+IR form the period also moves the number of
    refinement iterations (a whole solve sweep each), a jump that these
    features do not predict, so that form is tuned by measure().
"""
import math

import numpy as np

from . import FP64, FP32_IR, MLP_HIDDEN, MLP_INPUTS, MLP_PARAMS, mlp_predict, mlp_train

PERIODS = (1, 2, 3, 4)
MODEL_FORMS = (FP64,)  # forms the cost model is trained for and chooses in
AUTO = 0  # per level: 1 for levels whose longest front has < 48 steps, else 2
DEFAULT_PERIOD = 2


def largest_front(solver):
    """(n, c, w) of the analysed problem's front with the most pivot columns
    (the longest chain of 64-column steps; first of equal c)."""
    st = solver.structure()
    c, w = st["supernode_size"], st["work_size"]
    k = int(np.argmax(c))
    return int(c[k] + w[k]), int(c[k]), int(w[k])


def features(n, period, w, elem_bytes):
    """Model inputs: log2 order, update width in panels, log2(update rows + 1), bytes / 4."""
    return [math.log2(n), float(period), math.log2(w + 1), elem_bytes / 4.0]


def pick(times, prefer=DEFAULT_PERIOD, margin=0.005):
    """The fastest period of {period: ms}, but the preferred one unless the
    fastest beats it by more than margin (differences within run-to-run
    noise do not move the choice)."""
    best = min(times, key=times.get)
    if prefer in times and times[best] > times[prefer] * (1.0 - margin):
        return prefer
    return best


def measure(solver, a, precision, periods=PERIODS, reps=3, rhs=None):
    """{period: median factor ms} for an analysed solver (graph replays when
    the solver was created with graph=True; the first factorization of each
    period, which captures the graph, is not timed).  With rhs, the time is
    that of factor + solve: in the FP32+IR form the period changes the
    factor's rounding and with it, now and then, the number of refinement
    iterations.  Leaves the solver with the period pick() returns and a
    factorization made with it."""
    times = {}
    for p in periods:
        solver.update_period = p
        solver.factorize(a, precision)
        t = []
        for _ in range(reps + (rhs is not None)):
            solver.factorize(a, precision)
            ms = solver.stats()["factor_ms"]
            if rhs is not None:
                solver.solve(rhs)
                ms += solver.stats()["solve_ms"]
            t.append(ms)
        times[p] = float(np.median(t[-reps:]))
    best = pick(times)
    solver.update_period = best
    solver.factorize(a, precision)
    return times


def record(name, solver, a, rhs, precisions=(FP64, FP32_IR), periods=PERIODS, reps=3):
    """Tuner records of one analysed problem: dicts with the problem name,
    precision, period, front features and measured ms of factor + solve of
    rhs (the quantity the period is chosen for: in the FP32+IR form it also
    moves the number of refinement iterations)."""
    n, c, w = largest_front(solver)
    out = []
    for prec in precisions:
        eb = 8 if prec == FP64 else 4
        for p, ms in measure(solver, a, prec, periods, reps, rhs=rhs).items():
            out.append(dict(problem=name, precision=prec, period=p, n=n, c=c, w=w, elem_bytes=eb, ms=ms,
                            x=features(n, p, w, eb)))
    return out


class CostModel:
    """log(time) regression on the tuner's records, trained on the GPU."""

    def __init__(self, seed=0):
        rng = np.random.default_rng(seed)
        self.params = np.concatenate([
            rng.normal(0.0, 1.0 / math.sqrt(MLP_INPUTS), MLP_HIDDEN * MLP_INPUTS), np.zeros(MLP_HIDDEN),
            rng.normal(0.0, 1.0 / math.sqrt(MLP_HIDDEN), MLP_HIDDEN), [0.0]])
        assert len(self.params) == MLP_PARAMS
        self.mu = np.zeros(MLP_INPUTS)
        self.sd = np.ones(MLP_INPUTS)
        self.ty = (0.0, 1.0)
        self.loss = None

    def _x(self, x):
        return (np.asarray(x, dtype=np.float64) - self.mu) / self.sd

    def scale(self, records):
        """Sets the input and target scaling from records; returns their
        scaled inputs and targets (standardized log ms)."""
        x = np.array([r["x"] for r in records], dtype=np.float64)
        y = np.log(np.array([r["ms"] for r in records]))
        self.mu = x.mean(0)
        self.sd = np.where(x.std(0) > 0, x.std(0), 1.0)
        self.ty = (float(y.mean()), float(y.std()) or 1.0)
        return self._x(x), (y - self.ty[0]) / self.ty[1]

    def fit(self, records, epochs=4000, lr=0.01):
        x, t = self.scale(records)
        self.params, self.loss = mlp_train(self.params, x, t, epochs, lr)
        return self

    def predict_ms(self, x):
        y = mlp_predict(self.params, self._x(np.atleast_2d(x)))
        return np.exp(y * self.ty[1] + self.ty[0])

    def choose(self, n, w, elem_bytes, periods=PERIODS):
        """Period with the smallest predicted time for a front (n, w)."""
        pred = self.predict_ms([features(n, p, w, elem_bytes) for p in periods])
        return periods[int(np.argmin(pred))]


def regret(model, records, default=DEFAULT_PERIOD):
    """Per (problem, precision) group of records: the model's choice, its
    measured ms, the exhaustive tuner's best, regret = chosen / best - 1 and
    gain = default period's ms / chosen ms."""
    groups = {}
    for r in records:
        groups.setdefault((r["problem"], r["precision"]), []).append(r)
    out = []
    for key, rs in groups.items():
        t = {r["period"]: r["ms"] for r in rs}
        r0 = rs[0]
        p = model.choose(r0["n"], r0["w"], r0["elem_bytes"], tuple(sorted(t)))
        best = min(t.values())
        out.append(dict(problem=key[0], precision=key[1], chosen=p, best_period=min(t, key=t.get),
                        regret=t[p] / best - 1.0, gain=t.get(default, t[p]) / t[p]))
    return out


# --------------------------------------------------- per-front choice
# A front of order > 64 is factored either by 64-column steps spread over the
# whole GPU (large) or whole by one workgroup (medium, k_front_wg).  Medium
# fronts of one level run side by side, one per workgroup; large fronts of a
# level share each step launch.  The choice below is per front, made level by
# level without a sweep:
#   * FrontModel predicts a medium front's workgroup time from (order,
#     pivots, update rows, bytes per entry), trained by the mfd_mlp_* HIP
#     kernels on the per-front clocks the medium kernel itself records
#     (front_records);
#   * the large fronts' time comes from the level's step launches and Schur
#     launches, with constants measured once per form on a profiled
#     factorization (chain_constants): the time of a step launch per round of
#     diagonal tasks (one per CU), of a Schur launch, and the Schur rate;
#   * the model is trusted only inside the orders it was trained on (the
#     largest recorded front): larger fronts stay large;
#   * per level, the k largest fronts go to the steps and the rest to
#     workgroups, for the k that minimises the predicted level time.
# The exhaustive tuner (measure_split) times the factorization with each
# global pivot bound (fronts with at most cmax pivot columns medium) and is
# the yardstick of the policy's regret.
SPLITS = (0, 96, 128, 192, 256, 384, 1 << 30)
CUS = 256


APART_DIAG = 512  # the schedule's kApartDiag


def diag_slots(elem_bytes, ntasks=0):
    """Diagonal tasks of one step launch that run at once: one workgroup
    per CU at the FP64 step kernel's LDS, two at the FP32 one's; a step
    with at least APART_DIAG diagonal tasks runs them in a launch of their
    own (k_diag: two per CU in FP64, four in FP32, at its registers and
    LDS)."""
    if ntasks >= APART_DIAG:
        return CUS * (2 if elem_bytes == 8 else 4)
    return CUS * (1 if elem_bytes == 8 else 2)


def level_of(parent, rule=1):
    """Level of each node (0-based postorder, parent 1-based, 0 for roots),
    as the schedule builds its levels (mfd_set_level_rule): rule 1, one
    below the parent with the roots on the top level; rule 0, 1 + the
    highest child level."""
    m = len(parent)
    lev = np.zeros(m, dtype=np.int64)
    if rule == 1:
        depth = np.zeros(m, dtype=np.int64)
        for k in range(m - 1, -1, -1):  # parents follow their children
            p = int(parent[k]) - 1
            if p >= 0:
                depth[k] = depth[p] + 1
        return (depth.max() if m else 0) - depth
    for k in range(m):
        p = int(parent[k]) - 1
        if p >= 0:
            lev[p] = max(lev[p], lev[k] + 1)
    return lev


def measure_split(solver, a, precision, splits=SPLITS, reps=3):
    """{cmax: median factor ms} over the global pivot bounds (graph replays;
    the first factorization of each, which captures, is not timed).  Leaves
    the solver with no medium fronts."""
    times = {}
    for cmax in splits:
        solver.set_front_kinds(None)
        solver.set_front_split(cmax)
        solver.factorize(a, precision)
        t = []
        for _ in range(reps):
            solver.factorize(a, precision)
            t.append(solver.stats()["factor_ms"])
        times[cmax] = float(np.median(t))
    solver.set_front_split(0)
    return times


def measure_kinds(solver, a, precision, kinds, reps=3):
    """Median factor ms with the given per-node kinds."""
    solver.set_front_split(0)
    solver.set_front_kinds(kinds)
    solver.factorize(a, precision)
    t = []
    for _ in range(reps):
        solver.factorize(a, precision)
        t.append(solver.stats()["factor_ms"])
    solver.set_front_kinds(None)
    return float(np.median(t))


def front_features(n, c, w, elem_bytes):
    return np.stack([np.log2(n), np.log2(c), np.log2(np.asarray(w) + 1.0),
                     np.full(len(n), elem_bytes / 4.0)], axis=1).astype(np.float64)


def front_records(solver, a, precision, reps=2):
    """Per-front timings of the medium kernel: the analysed problem factored
    with every front of order > 64 medium; per front (n, c, w, bytes per
    entry) and its workgroup's microseconds (median of reps
    factorizations).  Leaves the solver with no medium fronts."""
    st = solver.structure()
    c, w = st["supernode_size"], st["work_size"]
    n = c + w
    solver.set_front_kinds(None)
    solver.set_front_split(1 << 30)
    solver.factorize(a, precision)
    ts = []
    for _ in range(reps):
        solver.factorize(a, precision)
        ts.append(solver.front_times())
    solver.set_front_split(0)
    t = np.median(np.stack(ts), axis=0)
    sel = t > 0
    eb = 8 if precision == FP64 else 4
    return dict(n=n[sel], c=c[sel], w=w[sel], elem_bytes=eb, us=t[sel])


class FrontModel:
    """log(microseconds) of a medium front's workgroup from (log2 order,
    log2 pivots, log2(update rows + 1), bytes / 4); the perceptron of the
    mfd_mlp_* kernels, trained on the GPU."""

    def __init__(self, seed=0):
        self.net = CostModel(seed)

    def fit(self, recs, epochs=4000, lr=0.01, max_samples=20000, seed=0):
        x = np.concatenate([front_features(r["n"], r["c"], r["w"], r["elem_bytes"]) for r in recs])
        y = np.concatenate([r["us"] for r in recs])
        if len(y) > max_samples:  # an even subsample of the fronts
            idx = np.random.default_rng(seed).choice(len(y), max_samples, replace=False)
            x, y = x[idx], y[idx]
        self.net.fit([dict(x=list(xi), ms=float(yi)) for xi, yi in zip(x, y)], epochs, lr)
        return self

    @property
    def loss(self):
        return self.net.loss

    def predict_us(self, n, c, w, elem_bytes):
        return self.net.predict_ms(front_features(np.asarray(n), np.asarray(c), np.asarray(w), elem_bytes))


def chain_record(ccs_col, ccs_row, max_zero, perm, a, precision, nb=64, batch=None):
    """Timings of the large-front schedule of one problem from a profiled
    factorization (every launch timed; no medium fronts): rounds of
    diagonal tasks (diag_slots at once) over all step launches and their
    microseconds, Schur launches, flops and microseconds.  batch: a list of
    (ccs_col, ccs_row) systems analysed together instead (perm: their
    orderings)."""
    from . import Solver
    s = Solver(nb=nb, profile=True)
    if batch is not None:
        s.analyze_batch(batch, max_zero, perm)
    else:
        s.analyze(ccs_col, ccs_row, max_zero, perm)
    s.factorize(a, precision)
    kt = s.kernel_times()
    st = s.structure()
    c, w, par = st["supernode_size"], st["work_size"], st["parent"]
    rule = s.level_rule
    s.close()
    n = c + w
    lev = level_of(par, rule)
    eb = 8 if precision == FP64 else 4
    rounds, flops = 0, 0.0
    for L in range(lev.max() + 1):
        big = (lev == L) & (n > 64)
        if not big.any():
            continue
        blocks = (c[big] + 63) // 64
        for t in range(-1, int(blocks.max()) - 1):
            nt = int((blocks > t + 1).sum())
            rounds += max(1, -(-nt // diag_slots(eb, nt)))
        flops += float((w[big].astype(np.float64) ** 2 * c[big]).sum())
    return dict(rounds=rounds, step_us=(kt["potrf"][0] + kt["trsm"][0]) * 1e3, schur_launches=kt["schur"][1],
                schur_flops=flops, schur_us=kt["schur"][0] * 1e3)


def chain_constants(records):
    """(microseconds per round of diagonal tasks, microseconds per Schur
    launch, Schur flops per microsecond) by least squares over chain
    records: step time = rounds * tau, Schur time = launches * t0 +
    flops / rate."""
    r = np.array([x["rounds"] for x in records], dtype=np.float64)
    tau = float((r * np.array([x["step_us"] for x in records])).sum() / (r * r).sum())
    M = np.array([[x["schur_launches"], x["schur_flops"]] for x in records], dtype=np.float64)
    y = np.array([x["schur_us"] for x in records])
    coef, *_ = np.linalg.lstsq(M, y, rcond=None)
    t0, inv_rate = max(float(coef[0]), 0.0), max(float(coef[1]), 1e-12)
    return tau, t0, 1.0 / inv_rate


def choose_kinds(structure, model, tau, schur_us, rate, elem_bytes, max_order, slots=2 * CUS, margin=0.1,
                 level_rule=1):
    """Per node: 1 (medium) or 0 (large), level by level, the k largest
    fronts (by pivots) large for the k that minimises the predicted level
    time: medium part max(longest front, sum / slots) by the model, large
    part (rounds of diagonal tasks) * tau + Schur launches (one per child
    rank) * schur_us + Schur flops / rate.  Fronts of order above max_order
    (outside the model's training data) stay large, and a level keeps all
    its fronts large unless the best split is predicted faster by more than
    margin (the predictions' error is of that order).  level_rule: the
    solver's (mfd_set_level_rule)."""
    c, w, par = structure["supernode_size"], structure["work_size"], structure["parent"]
    n = c + w
    lev = level_of(par, level_rule)
    kinds = np.zeros(len(c), dtype=np.int32)
    big = n > 64
    pred = np.zeros(len(c))
    if big.any():
        pred[big] = model.predict_us(n[big], c[big], w[big], elem_bytes)
    for L in range(lev.max() + 1):
        idx = np.flatnonzero((lev == L) & big)
        if len(idx) == 0:
            continue
        idx = idx[np.argsort(-c[idx], kind="stable")]
        m = len(idx)
        k_min = int((n[idx] > max_order).sum())  # untrusted: large
        idx = np.concatenate([idx[n[idx] > max_order], idx[n[idx] <= max_order]])
        idx = np.concatenate([idx[:k_min], idx[k_min:][np.argsort(-c[idx[k_min:]], kind="stable")]])
        blocks = (c[idx] + 63) // 64
        flops = np.concatenate([[0.0], np.cumsum(w[idx].astype(np.float64) ** 2 * c[idx])])
        p = pred[idx]
        suf_sum = np.concatenate([np.cumsum(p[::-1])[::-1], [0.0]])
        suf_max = np.concatenate([np.maximum.accumulate(p[::-1])[::-1], [0.0]])
        cands = sorted({k_min, m} | {min(m, k_min + int(round(1.5 ** e))) for e in range(0, 60) if 1.5 ** e <= 2 * m})
        best, best_k, t_large = None, 0, None
        for k in cands:
            t_med = max(suf_max[k], suf_sum[k] / slots) if k < m else 0.0
            t_chain = 0.0
            if k:
                b = blocks[:k]
                rounds = 0
                for t in range(-1, int(b.max()) - 1):
                    nt = int((b > t + 1).sum())
                    rounds += max(1, -(-nt // diag_slots(elem_bytes, nt)))
                t_chain = rounds * tau + schur_us + flops[k] / rate
            if k == m:
                t_large = t_med + t_chain
            if best is None or t_med + t_chain < best:
                best, best_k = t_med + t_chain, k
        if best >= (1.0 - margin) * t_large:
            best_k = m
        kinds[idx[best_k:]] = 1
    return kinds
