"""Float64 numpy restatement of the sizing MLP (16 -> 64 -> 64 -> 2, ReLU,
mean squared error), the CPU checker for the GPU forward / backward / Adam
kernels.  Parameter layout: W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]."""
import numpy as np

IN, H, OUT = 16, 64, 2
SIZES = [("W1", IN * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * OUT), ("b3", OUT)]
N_PARAMS = sum(n for _, n in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, n in SIZES:
        out[name] = p[o:o + n]
        o += n
    out["W1"] = out["W1"].reshape(IN, H)
    out["W2"] = out["W2"].reshape(H, H)
    out["W3"] = out["W3"].reshape(H, OUT)
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def init_params(seed: int = 0):
    """He-initialised weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2 / IN), (IN, H)), "b1": np.zeros(H),
         "W2": rng.normal(0, np.sqrt(2 / H), (H, H)), "b2": np.zeros(H),
         "W3": rng.normal(0, np.sqrt(1 / H), (H, OUT)), "b3": np.zeros(OUT)}
    return pack(d).astype(np.float32)


def forward(X, p):
    d = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ d["W1"] + d["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ d["W2"] + d["b2"]
    h2 = np.maximum(a2, 0)
    return h2 @ d["W3"] + d["b3"], (X, a1, h1, a2, h2)


def loss_and_grad(X, T, p):
    """Mean over rows of sum over outputs of (y - t)^2, and its gradient."""
    d = unpack(p)
    Y, (X, a1, h1, a2, h2) = forward(X, p)
    n = len(X)
    E = Y - np.asarray(T, dtype=np.float64)
    loss = float((E * E).sum() / n)
    dY = 2 * E / n
    g = {"W3": h2.T @ dY, "b3": dY.sum(0)}
    dh2 = (dY @ d["W3"].T) * (a2 > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ d["W2"].T) * (a1 > 0)
    g["W1"] = X.T @ dh1
    g["b1"] = dh1.sum(0)
    return loss, pack(g)


def adam(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    p = p - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return p, m1, m2
