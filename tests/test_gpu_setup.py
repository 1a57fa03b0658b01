"""The library's own set-up (modified Kerr-Schild geometry + Fishbone-Moncrief
torus, init.fishmon.c:237 and initbase.c:5 init) on the GPU against the
reference's initial state, then the reference's steps from it: the whole
pipeline with no reference data uploaded except the parameters."""
import numpy as np
import pytest

from golden_util import load

pytestmark = pytest.mark.gpu


def own_grid(d):
    from harm_finalwham_amd import harmgpu
    par = d["params"]
    n1, n2, n3, nb = (int(x) for x in par[:4])
    weno = int(d["weno"])
    g = harmgpu.Grid(n1, n2, n3, nb=nb, lim=harmgpu.HG_LIM_WENO5BND if weno else harmgpu.HG_LIM_MC,
                     fluxmethod=harmgpu.HG_FLUX_HLL if weno else harmgpu.HG_FLUX_LAXF, gam=par[5],
                     cour=par[10])
    g.build_mks_geometry(a=par[4], Rin=par[6], Rout=par[7], R0=par[9], hslope=par[8])
    dt = g.init_torus(rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7)
    return g, dt


@pytest.mark.parametrize("name", ["t2d_mc", "t2d_weno", "t3d_weno"])
def test_torus_initial_state_matches_reference(name):
    d = load(name)
    print("case", name, flush=True)
    g, dt = own_grid(d)
    assert np.array_equal(list(g.cfg.dx)[1:], d["params"][12:15])
    assert dt == d["params"][11]
    assert np.array_equal(g.get_prims(), d["p_init"])


@pytest.mark.parametrize("name", ["t2d_mc", "t3d_weno"])
def test_steps_from_own_initial_state_match_reference(name):
    d = load(name)
    print("case", name, flush=True)
    g, dt = own_grid(d)
    for n in range(int(d["nsteps"])):
        dt, nfail, _ = g.step_ch(dt)
        assert nfail == d["fail_seq"][n + 1]
        assert dt == d["dt_seq"][n + 1], n
    assert np.array_equal(g.get_prims(), d["p_final"])


def test_init_torus_needs_mks_geometry():
    from harm_finalwham_amd import harmgpu
    g = harmgpu.Grid(16, 16, 1, nb=3, lim=harmgpu.HG_LIM_MC, fluxmethod=harmgpu.HG_FLUX_LAXF)
    with pytest.raises(harmgpu.HarmError):
        g.init_torus()
