"""Samplers of the batched-inversion distribution on the host (numpy): the
Kerr-Schild metric (metric.c:697 set_gcov_ksmetric, with its closed-form
inverse) at random (r, theta) and random primitives.  Used by the fixture
generator (tests/golden/make_golden.py) and by the inversion tests for
subsets of the distribution (the tail of high Lorentz factor and
magnetisation)."""
import numpy as np


def ks_metric(r, th, a):
    """Kerr-Schild metric in (t, r, theta, phi) at BL radius r, theta
    (metric.c:697 set_gcov_ksmetric) and its closed-form inverse."""
    n = r.shape[0]
    s, c = np.sin(th), np.cos(th)
    s2 = s * s
    rho2 = r * r + a * a * c * c
    gcov = np.zeros((n, 4, 4))
    gcov[:, 0, 0] = -1.0 + 2.0 * r / rho2
    gcov[:, 0, 1] = gcov[:, 1, 0] = 2.0 * r / rho2
    gcov[:, 0, 3] = gcov[:, 3, 0] = -2.0 * a * r * s2 / rho2
    gcov[:, 1, 1] = 1.0 + 2.0 * r / rho2
    gcov[:, 1, 3] = gcov[:, 3, 1] = -a * s2 * (1.0 + 2.0 * r / rho2)
    gcov[:, 2, 2] = rho2
    gcov[:, 3, 3] = s2 * (rho2 + a * a * s2 * (1.0 + 2.0 * r / rho2))
    delta = r * r - 2.0 * r + a * a
    gcon = np.zeros((n, 4, 4))
    gcon[:, 0, 0] = -(1.0 + 2.0 * r / rho2)
    gcon[:, 0, 1] = gcon[:, 1, 0] = 2.0 * r / rho2
    gcon[:, 1, 1] = delta / rho2
    gcon[:, 1, 3] = gcon[:, 3, 1] = a / rho2
    gcon[:, 2, 2] = 1.0 / rho2
    gcon[:, 3, 3] = 1.0 / (rho2 * s2)
    gdet = rho2 * s
    gpert = np.zeros((n, 4))
    gpert[:, 0] = 2.0 * r / rho2
    gpert[:, 1] = 2.0 * r / rho2
    gpert[:, 2] = gcov[:, 2, 2] - 1.0
    gpert[:, 3] = gcov[:, 3, 3] - 1.0
    return gcov, gcon, gdet, gpert


def sample_c2p(n, rng, a=0.9375, lor=(1.0, 50.0), sigma=(1e-4, 1e3)):
    """The batched-inversion distribution: rho log-U[1e-6,1], u log-U[1e-8,1e2],
    Lorentz factor log-U[lor] (default [1,50]) with isotropic direction,
    b^2/rho log-U[sigma] (default [1e-4,1e3]) with isotropic direction, KS
    metric at r U[1.3,40], theta U(0.01,pi-0.01).  Returns gcov, gcon [n,4,4],
    gdet [n], gcovpert [n,4], primitives [n,8] and a guess [n,8] (u and the
    velocities perturbed by 1e-2 N(0,1))."""
    r = rng.uniform(1.3, 40.0, n)
    th = rng.uniform(0.01, np.pi - 0.01, n)
    gcov, gcon, gdet, gpert = ks_metric(r, th, a)
    rho = 10.0 ** rng.uniform(-6, 0, n)
    u = 10.0 ** rng.uniform(-8, 2, n)
    lor = 10.0 ** rng.uniform(np.log10(lor[0]), np.log10(lor[1]), n)
    sig = 10.0 ** rng.uniform(np.log10(sigma[0]), np.log10(sigma[1]), n)
    gs = gcov[:, 1:, 1:]

    def unit(v):
        nrm = np.sqrt(np.einsum("ni,nij,nj->n", v, gs, v))
        return v / nrm[:, None]

    ev = unit(rng.normal(size=(n, 3)))
    ut = np.sqrt(lor * lor - 1.0)[:, None] * ev
    eb = unit(rng.normal(size=(n, 3)))
    # b^2 for unit-amplitude B (phys.c:849 get_state), then rescale to b^2/rho = sig
    alpha = 1.0 / np.sqrt(-gcon[:, 0, 0])
    beta = gcon[:, 0, 1:] * (alpha * alpha)[:, None]
    gam_n = np.sqrt(1.0 + np.einsum("ni,nij,nj->n", ut, gs, ut))
    ucon = np.zeros((n, 4))
    ucon[:, 0] = gam_n / alpha
    ucon[:, 1:] = ut - ucon[:, :1] * beta
    ucov = np.einsum("nij,nj->ni", gcov, ucon)
    bt = np.einsum("ni,ni->n", eb, ucov[:, 1:])
    bcon = np.zeros((n, 4))
    bcon[:, 0] = bt
    bcon[:, 1:] = (eb + bt[:, None] * ucon[:, 1:]) / ucon[:, :1]
    bsq1 = np.einsum("ni,nij,nj->n", bcon, gcov, bcon)
    amp = np.sqrt(sig * rho / bsq1)
    p = np.zeros((n, 8))
    p[:, 0] = rho
    p[:, 1] = u
    p[:, 2:5] = ut
    p[:, 5:8] = amp[:, None] * eb
    guess = p.copy()
    guess[:, 1] *= 1.0 + 1e-2 * rng.normal(size=n)
    guess[:, 2:5] *= 1.0 + 1e-2 * rng.normal(size=(n, 3))
    return gcov, gcon, gdet, gpert, p, guess
