"""CPU FP64 reference of the learned-guess perceptron (test infrastructure).

The network is the one the library runs (8 -> 64 -> 64 -> 1, tanh), with
the same parameter block layout (W1[8][64] b1[64] W2[64][64] b2[64] W3[64]
b3 mu[8] sd[8]).  With quant=True it applies the library's number formats:
hidden-layer weights rounded to BF16 (round to nearest even), the network
input carried as a hi + lo BF16 pair (hi = bf16(x), lo = bf16(x - hi)), the
first layer's activations rounded to BF16 before the second layer; the
gradient treats these roundings as the identity, as the library does.  With
quant=False it is the plain FP64 network, which the finite-difference check
pins.
"""
import numpy as np

NIN, NH = 8, 64
OFF = {"W1": 0, "b1": 512, "W2": 576, "b2": 4672, "W3": 4736, "b3": 4800, "mu": 4801, "sd": 4809}
NPARAM = 4817
NTRAIN = OFF["mu"]


def bf16(x):
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b >> 16) & 1) + 0x7FFF
    return ((b + r) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def hilo(v):
    """The value a layer input carries into the matrix cores: bf16(v) + bf16(v - bf16(v))."""
    v = np.asarray(v, np.float32)
    hi = bf16(v)
    lo = bf16(v - hi)
    return (hi + lo).astype(np.float32)


def unpack(w):
    w = np.asarray(w, np.float64)
    return dict(W1=w[OFF["W1"]:OFF["b1"]].reshape(NIN, NH), b1=w[OFF["b1"]:OFF["W2"]],
                W2=w[OFF["W2"]:OFF["b2"]].reshape(NH, NH), b2=w[OFF["b2"]:OFF["W3"]], W3=w[OFF["W3"]:OFF["b3"]],
                b3=w[OFF["b3"]], mu=w[OFF["mu"]:OFF["sd"]], sd=w[OFF["sd"]:])


def forward(w, X, quant=True):
    """Returns (y, cache) for raw features X [n, 8]."""
    p = unpack(w)
    if quant:
        w32 = np.asarray(w, np.float32)
        x = (np.asarray(X, np.float32) - w32[OFF["mu"]:OFF["sd"]]) / w32[OFF["sd"]:]
        x = hilo(x).astype(np.float64)
        W1 = bf16(w32[OFF["W1"]:OFF["b1"]]).astype(np.float64).reshape(NIN, NH)
        W2 = bf16(w32[OFF["W2"]:OFF["b2"]]).astype(np.float64).reshape(NH, NH)
    else:
        x = (np.asarray(X, np.float64) - p["mu"]) / p["sd"]
        W1, W2 = p["W1"], p["W2"]
    h1 = np.tanh(x @ W1 + p["b1"])
    h1in = bf16(h1.astype(np.float32)).astype(np.float64) if quant else h1
    h2 = np.tanh(h1in @ W2 + p["b2"])
    y = h2 @ p["W3"] + p["b3"]
    return y, dict(x=x, h1=h1, h1in=h1in, h2=h2, W2=W2)


def loss_grad(w, X, Y, quant=True):
    """Mean squared error and its gradient w.r.t. the parameter block
    (normalisation entries 0)."""
    p = unpack(w)
    y, c = forward(w, X, quant)
    n = len(Y)
    r = y - np.asarray(Y, np.float64)
    loss = np.mean(r * r)
    dy = 2.0 * r / n
    g = np.zeros(NPARAM)
    g[OFF["W3"]:OFF["b3"]] = c["h2"].T @ dy
    g[OFF["b3"]] = dy.sum()
    dz2 = dy[:, None] * p["W3"][None, :] * (1.0 - c["h2"] ** 2)
    g[OFF["W2"]:OFF["b2"]] = (c["h1in"].T @ dz2).ravel()
    g[OFF["b2"]:OFF["W3"]] = dz2.sum(0)
    dz1 = (dz2 @ c["W2"].T) * (1.0 - c["h1"] ** 2)
    g[OFF["W1"]:OFF["b1"]] = (c["x"].T @ dz1).ravel()
    g[OFF["b1"]:OFF["W2"]] = dz1.sum(0)
    return loss, g


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM)
    w[:NTRAIN] = rng.normal(0.0, 0.3 * scale, NTRAIN)
    w[OFF["mu"]:OFF["sd"]] = rng.normal(0.0, 1.0, NIN)
    w[OFF["sd"]:] = rng.uniform(0.5, 2.0, NIN)
    return w
