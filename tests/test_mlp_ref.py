"""The CPU reference of the learned-guess perceptron against mathematics:
central finite differences of its loss pin the backpropagation, and the
BF16 rounding helper against exactly representable and tie cases."""
import numpy as np

import mlp_ref


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = mlp_ref.random_params(5)
    X = rng.normal(0.0, 1.5, (37, mlp_ref.NIN))
    Y = rng.normal(0.0, 0.5, 37)
    _, g = mlp_ref.loss_grad(w, X, Y, quant=False)
    # every bias and W3 entry, and a sample of W1 and W2 entries
    idx = list(range(mlp_ref.OFF["b1"], mlp_ref.OFF["W2"])) + list(range(mlp_ref.OFF["b2"], mlp_ref.NTRAIN))
    idx += list(rng.choice(np.arange(mlp_ref.OFF["W1"], mlp_ref.OFF["b1"]), 64, replace=False))
    idx += list(rng.choice(np.arange(mlp_ref.OFF["W2"], mlp_ref.OFF["b2"]), 256, replace=False))
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (mlp_ref.loss_grad(wp, X, Y, quant=False)[0] - mlp_ref.loss_grad(wm, X, Y, quant=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(fd)), (k, fd, g[k])
    # the normalisation entries are not trained
    assert np.all(g[mlp_ref.NTRAIN:] == 0.0)


def test_bf16_rounding():
    x = np.array([1.0, -2.5, 0.0, 1.0 + 2.0 ** -8, 1.0 + 2.0 ** -7, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9],
                 np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins;
    # 1 + 3 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    want = np.array([1.0, -2.5, 0.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, 1.0], np.float32)
    np.testing.assert_array_equal(mlp_ref.bf16(x), want)
    v = np.float32(0.123456789)
    # hi + lo keeps at least 16 significant bits
    assert abs(float(mlp_ref.hilo(v)) - float(v)) <= 2.0 ** -16 * abs(float(v))


def test_quantised_network_close_to_plain():
    rng = np.random.default_rng(9)
    w = mlp_ref.random_params(2)
    X = rng.normal(0.0, 1.0, (200, mlp_ref.NIN))
    yq, _ = mlp_ref.forward(w, X, quant=True)
    yp, _ = mlp_ref.forward(w, X, quant=False)
    # only the weight rounding (relative 2^-9) separates them
    assert np.max(np.abs(yq - yp)) < 0.05 * np.max(np.abs(yp))
    assert np.max(np.abs(yq - yp)) > 0.0
