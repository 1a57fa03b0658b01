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


def state_digest(p):
    """SHA-256 of a state array with -0.0 folded into +0.0 (the digest the
    golden generator stores for states too large to keep)."""
    import hashlib
    a = np.ascontiguousarray(p, dtype="<f8") + 0.0
    return hashlib.sha256(a.tobytes()).hexdigest()


def divb_fluxct(p, gdet_cent, gdet_corn, dx, nb):
    """Normalised flux-CT divergence of B at the interior zone corners
    (global.h:1586-1645, DIVBFLUXCT with DIVBNORMFLUXCT): p[i][j][k][pl]
    with ghosts, gdet_cent / gdet_corn [i][j] at zone centres / corners on
    the same index range.  Returns |div B| normalised by the local field."""
    n1, n2, n3 = (s - 2 * nb if s > 1 else 1 for s in p.shape[:3])
    s3 = p.shape[2] > 1
    g = gdet_cent[:, :, None]
    Fg = [p[..., 5 + d] * g for d in range(3)]

    def sh(a, axis):  # value at index - 1 along axis (identity along a flat axis)
        if axis == 2 and not s3:
            return a
        return np.roll(a, 1, axis=axis)

    def avg(a, axis):
        return 0.5 * (a + sh(a, axis))

    A1 = 0.5 * (avg(Fg[0], 1) + sh(avg(Fg[0], 1), 2))
    A2 = 0.5 * (avg(Fg[1], 0) + sh(avg(Fg[1], 0), 2))
    A3 = 0.5 * (avg(Fg[2], 1) + sh(avg(Fg[2], 1), 0))
    dif = (A1 - sh(A1, 0)) / dx[1] + (A2 - sh(A2, 1)) / dx[2]
    nrm = (A1 + sh(A1, 0)) + (A2 + sh(A2, 1))
    if s3:
        dif = dif + (A3 - sh(A3, 2)) / dx[3]
        nrm = nrm + (A3 + sh(A3, 2))
    else:
        dif = dif + (A3 - A3) / dx[3]
        nrm = nrm + (A3 + A3)
    div = np.abs(dif) * max(dx[1], dx[2], dx[3]) / ((1.0 / 3.0) * gdet_corn[:, :, None] * np.abs(nrm) + 1e-50)
    # interior corners only: i in [1, n1-1], j in [1, n2-1], k in [1, n3-1]
    ks = slice(nb + 1, nb + n3) if s3 else slice(None)
    return div[nb + 1:nb + n1, nb + 1:nb + n2, ks]
