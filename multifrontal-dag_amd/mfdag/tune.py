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
