"""numpy restatement of the schedule tuner's cost model (the HIP kernels in
multifrontal-dag_amd/csrc/mfd_mlp.hip).

TEST INFRASTRUCTURE ONLY: only tests/ may import it, as the checker.

The model has no counterpart in the reference (whose schedule is StarPU's),
so nothing in the reference pins it.  It is pinned instead by what the
mathematics fixes (tests/test_tuning.py): the gradient against central
differences of the loss, the closed forms of a network with zero output
weights (output b2, gradient 2 (b2 - mean t) on b2 and zero on W1, b1), and
Adam's first step (each parameter moves by lr sign(g) to within eps).

Parameters (193): W1 (32 x 4, row-major), b1 (32), w2 (32), b2.
y(x) = b2 + w2 . tanh(W1 x + b1); loss = mean (y - t)^2.
"""
import numpy as np

N_IN, N_HID = 4, 32
N_PAR = N_HID * (N_IN + 2) + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[:N_HID * N_IN].reshape(N_HID, N_IN)
    b1 = p[N_HID * N_IN:N_HID * N_IN + N_HID]
    w2 = p[N_HID * N_IN + N_HID:N_HID * N_IN + 2 * N_HID]
    return W1, b1, w2, p[-1]


def predict(p, x):
    W1, b1, w2, b2 = unpack(p)
    return np.tanh(np.asarray(x, dtype=np.float64) @ W1.T + b1) @ w2 + b2


def loss(p, x, t):
    return float(np.mean((predict(p, x) - np.asarray(t)) ** 2))


def gradient(p, x, t):
    """(gradient, loss) of the mean squared error, by back-propagation."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    W1, b1, w2, b2 = unpack(p)
    h = np.tanh(x @ W1.T + b1)
    r = h @ w2 + b2 - t
    n = len(x)
    e = 2.0 * r / n
    g = np.outer(e, w2) * (1.0 - h * h)
    grad = np.concatenate([(g.T @ x).ravel(), g.sum(0), h.T @ e, [e.sum()]])
    return grad, float(np.sum(r * r) / n)


def train(p, x, t, epochs, lr):
    """Full-batch Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected),
    moments from zero; returns (params, loss before the last step)."""
    p = np.array(p, dtype=np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    last = None
    for it in range(1, epochs + 1):
        g, last = gradient(p, x, t)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        mh = m / (1.0 - 0.9 ** it)
        vh = v / (1.0 - 0.999 ** it)
        p -= lr * mh / (np.sqrt(vh) + 1e-8)
    return p, last
