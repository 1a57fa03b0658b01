"""Helpers shared by the tests: fixture loading and small comparisons."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def active(a, nb, n3):
    """Active zones of a [i][j][k][pl] array with nb ghost layers (k ghosts only in 3-D)."""
    nb3 = nb if n3 > 1 else 0
    return a[nb:-nb, nb:-nb, nb3:a.shape[2] - nb3]


def max_rel(a, b, floor=1e-300):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
