"""Longer and larger runs of the library against the reference:
  the 128x128 MC/LaxF torus over 100 RK2 steps, from the library's own
  initial data: every dt, no failed inversion, the final state bit for bit,
  div B at round-off;
  the 256x128x64 WENO5BND/HLL torus (the benchmark grid) over one RK2 step
  and over 50 (every dt and failure count, the final state's digest, and
  the reference's constotal and divbmaxavg of it), and the same grid in the
  torus file's own scheme (PARA + HLL + RK4, finite volume) over one step:
  digests of the initial and final states and the dt sequence; div B at
  round-off over 100 steps of both schemes on the benchmark grid."""
import os

import numpy as np
import pytest

from golden_util import GOLDEN, divb_fluxct, load, state_digest

pytestmark = pytest.mark.gpu


def _own(d, nb):
    from harm_finalwham_amd import harmgpu
    from test_gpu_step import scheme
    par = d["params"]
    n1, n2, n3 = (int(x) for x in par[:3])
    lim, flux, timeorder, doenoflux = scheme(d)
    g = harmgpu.Grid(n1, n2, n3, nb=nb, lim=lim, fluxmethod=flux, gam=par[5], cour=par[10], timeorder=timeorder,
                     doenoflux=doenoflux)
    g.build_mks_geometry(a=par[4], Rin=par[6], Rout=par[7], R0=par[9], hslope=par[8])
    dt = g.init_torus(rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7)
    return g, dt


def test_config1_100_steps_match_reference():
    from harm_finalwham_amd import harmgpu
    d = load("c1_mc_100")
    nb = int(d["params"][3])
    print("case 128x128 MC/LaxF, 100 steps", flush=True)
    g, dt = _own(d, nb)
    assert dt == d["dt_seq"][0]
    assert state_digest(g.get_prims()) == str(d["digest_init"])
    for n in range(1, int(d["nsteps"]) + 1):
        dt, nfail, _ = g.step_ch(dt)
        assert nfail == d["fail_seq"][n]
        assert dt == d["dt_seq"][n], n
    p = g.get_prims()
    assert np.array_equal(p, d["p_final"])
    geo = harmgpu.mks_geometry_tables(128, 128, 1, nb)
    m = 128 + 2 * nb
    div = divb_fluxct(p, geo["gdet"][:m, :m, 6], geo["gdet"][:m, :m, 5], geo["dx"], nb)
    # the corner rows next to the axis carry POLEDEATH's zeroing of B2
    assert div[:, 1:-1].max() < 1e-11


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "f3d_weno_1.npz")),
                    reason="full-size fixture not generated (make_golden.py --long --full)")
def test_benchmark_grid_one_step_matches_reference():
    d = load("f3d_weno_1")
    nb = int(d["params"][3])
    print("case 256x128x64 WENO5BND/HLL, 1 step", flush=True)
    g, dt = _own(d, nb)
    assert dt == d["dt_seq"][0]
    assert state_digest(g.get_prims()) == str(d["digest_init"])
    dt, nfail, _ = g.step_ch(dt)
    assert nfail == d["fail_seq"][1]
    assert dt == d["dt_seq"][1]
    p = g.get_prims()
    assert np.array_equal(p.reshape(-1, 8).sum(axis=0), d["sum_final"])
    assert state_digest(p) == str(d["digest_final"])


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "f3d_weno_50.npz")),
                    reason="full-size fixture not generated (make_golden.py --long50)")
def test_benchmark_grid_50_steps_match_reference():
    """The benchmark grid at its configuration's length (step_ch.c:35, 50
    RK2 steps of 256x128x64 WENO5BND + HLL): every dt and failure count, the
    final state, constotal (dump_ener.c:596) and divbmaxavg
    (dump_ener.c:482) against the reference build's own run."""
    d = load("f3d_weno_50")
    par = load("f3d_weno_1")
    nb = int(par["params"][3])
    print("case 256x128x64 WENO5BND/HLL, 50 steps", flush=True)
    g, dt = _own(par, nb)
    assert dt == d["dt_seq"][0]
    assert state_digest(g.get_prims()) == str(d["digest_init"])
    for n in range(1, int(d["nsteps"]) + 1):
        dt, nfail, _ = g.step_ch(dt)
        assert nfail == d["fail_seq"][n], n
        assert dt == d["dt_seq"][n], n
    p = g.get_prims()
    assert np.array_equal(p.reshape(-1, 8).sum(axis=0), d["sum_final"])
    assert state_digest(p) == str(d["digest_final"])
    r = g.diagnostics(0.0)
    assert np.array_equal(r["cons"], d["cons"])
    # the normalisation of div B is long double in the reference
    assert abs(r["divbmax"] - d["divb"][0]) <= 1e-14 * abs(d["divb"][0])
    assert abs(r["divbavg"] - d["divb"][1]) <= 1e-13 * abs(d["divb"][1])
    g.close()


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "f3d_fv_1.npz")),
                    reason="full-size fixture not generated (make_golden.py --long --fv)")
def test_benchmark_grid_torus_scheme_one_step_matches_reference():
    """The torus file's own scheme (PARA + HLL + RK4, finite volume) on the
    benchmark grid: one step, digests of the whole state."""
    d = load("f3d_fv_1")
    nb = int(d["params"][3])
    print("case 256x128x64 PARA/HLL/RK4 finite volume, 1 step", flush=True)
    g, dt = _own(d, nb)
    assert dt == d["dt_seq"][0]
    assert state_digest(g.get_prims()) == str(d["digest_init"])
    dt, nfail, _ = g.step_ch(dt)
    if d["fail_seq"][1]:
        assert nfail > 0
    assert dt == d["dt_seq"][1]
    p = g.get_prims()
    assert np.array_equal(p.reshape(-1, 8).sum(axis=0), d["sum_final"])
    assert state_digest(p) == str(d["digest_final"])


@pytest.mark.parametrize("scheme", ["weno", "fv"])
def test_benchmark_grid_keeps_div_b_at_round_off(scheme):
    """Size-independent property at the benchmark size: flux-CT keeps the
    corner divergence of B at round-off over 100 steps of the 256x128x64
    torus (WENO5BND + HLL + RK2, and the torus file's finite-volume PARA +
    HLL + RK4), with no failed step."""
    from harm_finalwham_amd import harmgpu
    print("case 256x128x64 div B over 100 steps,", scheme, flush=True)
    nb = 6 if scheme == "weno" else 3
    if scheme == "weno":
        g = harmgpu.Grid(256, 128, 64, nb=nb, lim=harmgpu.HG_LIM_WENO5BND, fluxmethod=harmgpu.HG_FLUX_HLL)
    else:
        g = harmgpu.Grid(256, 128, 64, nb=nb, lim=harmgpu.HG_LIM_PARA, fluxmethod=harmgpu.HG_FLUX_HLL, timeorder=4,
                         doenoflux=harmgpu.HG_ENOFINITEVOLUME)
    g.build_mks_geometry()
    dt = g.init_torus(rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7)
    geo = harmgpu.mks_geometry_tables(256, 128, 64, nb)
    m1, m2 = 256 + 2 * nb, 128 + 2 * nb
    d0 = divb_fluxct(g.get_prims(), geo["gdet"][:m1, :m2, 6], geo["gdet"][:m1, :m2, 5], geo["dx"], nb)
    for n in range(100):
        dt, _, _ = g.step_ch(dt)
        assert np.isfinite(dt) and dt > 0
    p = g.get_prims()
    assert np.isfinite(p).all()
    d1 = divb_fluxct(p, geo["gdet"][:m1, :m2, 6], geo["gdet"][:m1, :m2, 5], geo["dx"], nb)
    # the corner rows next to the axis carry POLEDEATH's zeroing of B2
    print("max |div B| (normalised) initial %.3g, after 100 steps %.3g" % (d0[:, 1:-1].max(), d1[:, 1:-1].max()),
          flush=True)
    assert d1[:, 1:-1].max() < 1e-10
