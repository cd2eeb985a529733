"""TEST INFRASTRUCTURE, not product code: FP64 restatement of the absorbing-strip
network, used only by tests/ as the checker of the GPU training kernels.

The network has no counterpart in the reference. Definition: x in R^16,
h1 = tanh(W1 x + b1), h2 = tanh(W2 h1 + b2), y = W3 h2 + b3 (hidden 64, y in R^2),
one flat FP32 parameter vector W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
Damping head: g = GMAX * sigmoid(y); prediction pred = (-x_0 g_0, -x_7 g_1) (the
change of the cell's P/s and Q/s); loss = mean over samples of sum over outputs of
(pred - t)^2.

The gradient below is hand-written backpropagation; tests pin it against central
finite differences of `loss` (which only needs the forward pass).
"""
import numpy as np

NIN, NH, NOUT = 16, 64, 2
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NOUT, NH)), ("b3", (NOUT,))]
NPAR = sum(int(np.prod(s)) for _, s in SHAPES)
GMAX = 0.25
HEAD_IN = (0, 7)


def unpack(p):
    p = np.asarray(p, dtype=np.float64).ravel()
    if p.size != NPAR:
        raise ValueError(f"{p.size} parameters, expected {NPAR}")
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[k:k + n].reshape(shape)
        k += n
    return out


def init(seed, scale=1.0, out_bias=-4.0):
    """Seeded initial parameters: Glorot-uniform weights, zero biases, output layer times 0.1,
    output bias out_bias (a nearly closed damping head)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1])) * scale * (0.1 if name == "W3" else 1.0)
            parts.append(rng.uniform(-lim, lim, size=shape).ravel())
        else:
            parts.append(np.full(shape, out_bias if name == "b3" else 0.0).ravel())
    return np.concatenate(parts).astype(np.float32)


def raw(p, X):
    """Raw network outputs y."""
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.tanh(X @ w["W1"].T + w["b1"])
    h2 = np.tanh(h1 @ w["W2"].T + w["b2"])
    y = h2 @ w["W3"].T + w["b3"]
    return y, (X, h1, h2)


def forward(p, X):
    """Damping-head prediction pred (n, 2) and the intermediates."""
    y, (X, h1, h2) = raw(p, X)
    g = GMAX / (1.0 + np.exp(-y))
    xs = X[:, list(HEAD_IN)]
    return -xs * g, (X, h1, h2, g, xs)


MMAX = 3.0


def forward_coef(p, X):
    """Coefficient-head training target: pred = (m, 0), m = MMAX * sigmoid(y_0) (output 1 unused)."""
    y, (X, h1, h2) = raw(p, X)
    m = MMAX / (1.0 + np.exp(-y[:, 0]))
    return np.stack([m, np.zeros_like(m)], axis=1), (X, h1, h2, m)


def loss(p, X, T, head=0):
    """Mean over samples of the summed squared error; head 2: of output 0 (the multiplier) only."""
    T = np.asarray(T, np.float64)
    if head == 2:
        pred, _ = forward_coef(p, X)
        return float(np.sum((pred[:, 0] - T[:, 0]) ** 2) / len(pred))
    pred, _ = forward(p, X)
    return float(np.sum((pred - T) ** 2) / len(pred))


def loss_grad(p, X, T, head=0):
    """Loss and its gradient (flat, FP64) by backpropagation (head as in `loss`)."""
    w = unpack(p)
    T = np.asarray(T, np.float64)
    if head == 2:
        pred, (X, h1, h2, m) = forward_coef(p, X)
        n = len(pred)
        e = np.zeros_like(pred)
        e[:, 0] = pred[:, 0] - T[:, 0]
        dy = np.zeros_like(pred)
        dy[:, 0] = 2.0 * e[:, 0] / n * m * (1.0 - m / MMAX)
    else:
        pred, (X, h1, h2, g, xs) = forward(p, X)
        n = len(pred)
        e = pred - T
        dy = 2.0 * e / n * (-xs * g * (1.0 - g / GMAX))
    grad = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ w["W3"]) * (1.0 - h2 ** 2)
    grad["W2"] = d2.T @ h1
    grad["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"]) * (1.0 - h1 ** 2)
    grad["W1"] = d1.T @ X
    grad["b1"] = d1.sum(0)
    return float(np.sum(e ** 2) / n), np.concatenate([grad[name].ravel() for name, _ in SHAPES])
