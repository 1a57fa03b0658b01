"""Full torus steps on the GPU (ghost fill, reconstruction, fluxes,
flux-CT, update, Noble 2D, fixup1zone, post_fixup) against the reference
golden vectors: bit-for-bit primitives and time steps.  Schemes: MC + LaxF
and WENO5BND + HLL with the midpoint RK2, PARA + HLL with RK4, the last with
the finite-volume step (the torus file's own scheme), MC + LaxF with TVD RK3
and with Euler, the FORCE, MUSTA and HLLLAXF1 fluxes; and the reference's
step loop driven substep by substep through hg_advance_substep."""
import numpy as np
import pytest

from golden_util import load

pytestmark = pytest.mark.gpu

TORI = ["t2d_mc", "t2d_weno", "t3d_weno", "t2d_para", "t3d_para", "t2d_fv", "t3d_fv", "t2d_rk3", "t2d_euler",
        "t2d_force", "t2d_musta", "t2d_hlllaxf1", "t2d_fvweno", "t3d_fvweno", "t2d_fv_30", "t3d_weno_10", "t3d_fv_10",
        "t2d_vanl", "t2d_minm", "t2d_donor"]


def scheme(d):
    """(lim, fluxmethod, timeorder, doenoflux) of a torus fixture's scheme number."""
    from harm_finalwham_amd import harmgpu as h
    return {0: (h.HG_LIM_MC, h.HG_FLUX_LAXF, 2, h.HG_NOENOFLUX),
            1: (h.HG_LIM_WENO5BND, h.HG_FLUX_HLL, 2, h.HG_NOENOFLUX),
            2: (h.HG_LIM_PARA, h.HG_FLUX_HLL, 4, h.HG_NOENOFLUX),
            3: (h.HG_LIM_PARA, h.HG_FLUX_HLL, 4, h.HG_ENOFINITEVOLUME),
            4: (h.HG_LIM_MC, h.HG_FLUX_LAXF, 3, h.HG_NOENOFLUX),
            5: (h.HG_LIM_MC, h.HG_FLUX_LAXF, 1, h.HG_NOENOFLUX),
            6: (h.HG_LIM_MC, h.HG_FLUX_FORCE, 2, h.HG_NOENOFLUX),
            7: (h.HG_LIM_PARA, h.HG_FLUX_MUSTA, 2, h.HG_NOENOFLUX),
            8: (h.HG_LIM_WENO5BND, h.HG_FLUX_HLLLAXF1, 2, h.HG_NOENOFLUX),
            9: (h.HG_LIM_WENO5BND, h.HG_FLUX_HLL, 2, h.HG_ENOFINITEVOLUME),
            10: (h.HG_LIM_VANL, h.HG_FLUX_LAXF, 2, h.HG_NOENOFLUX),
            11: (h.HG_LIM_MINM, h.HG_FLUX_HLL, 2, h.HG_NOENOFLUX),
            12: (h.HG_LIM_DONOR, h.HG_FLUX_LAXF, 2, h.HG_NOENOFLUX)}[int(d["weno"])]


def make_grid(d):
    from harm_finalwham_amd import harmgpu
    par = d["params"]
    n1, n2, n3, nb = (int(x) for x in par[:4])
    lim, flux, timeorder, doenoflux = scheme(d)
    g = harmgpu.Grid(n1, n2, n3, nb=nb, lim=lim, fluxmethod=flux, gam=par[5], cour=par[10],
                     dx=(0.0, par[12], par[13], par[14]), timeorder=timeorder, doenoflux=doenoflux)
    g.upload_geometry(d["gcov"], d["gcon"], d["gcovpert"], d["gdet"], d["conn"], d["xv_cent"][:, :, 1, 1])
    return g


@pytest.mark.parametrize("name", ["t2d_mc", "t3d_weno"])
def test_bound_prim_reproduces_reference(name):
    d = load(name)
    g = make_grid(d)
    g.set_prims(d["p_step1"])
    g.bound_prim()
    assert np.array_equal(g.get_prims(), d["p_step1"])


@pytest.mark.parametrize("name", TORI)
def test_fluxcalc_matches_reference(name):
    d = load(name)
    g = make_grid(d)
    nb = int(d["params"][3])
    g.set_prims(d["p_init"])
    F, ndt = g.fluxcalc()
    assert np.array_equal(ndt, d["ndt_init"])
    n3 = int(d["params"][2])
    k = slice(nb, -nb) if n3 > 1 else slice(None)
    kf = slice(nb, -nb + 1) if n3 > 1 else slice(None)
    assert np.array_equal(F[0][nb:-nb + 1, nb:-nb, k], d["F1_init"][nb:-nb + 1, nb:-nb, k])
    assert np.array_equal(F[1][nb:-nb, nb:-nb + 1, k], d["F2_init"][nb:-nb, nb:-nb + 1, k])
    if n3 > 1:
        assert np.array_equal(F[2][nb:-nb, nb:-nb, kf], d["F3_init"][nb:-nb, nb:-nb, kf])


@pytest.mark.parametrize("name", TORI)
def test_steps_match_reference(name):
    import pyoracle
    d = load(name)
    g = make_grid(d)
    g.set_prims(d["p_init"])
    G = pyoracle.Grid.from_fixture(d)
    po = d["p_init"]
    Uo = G.cons_from_prims(po) if G.g.fv else None
    dt = float(d["dt_seq"][0])
    for n in range(1, int(d["nsteps"]) + 1):
        if Uo is None:
            po, _, nf_oracle, _ = G.step(po, dt)
        else:
            po, Uo, _, nf_oracle, _ = G.step_fv(po, Uo, dt)
        dt, nf, it = g.step_ch(dt)
        print(name, "step", n, "failed inversions", nf, flush=True)
        # failed inversions over the step's substeps, as the oracle counts them
        assert nf == nf_oracle and it > 0
        assert dt == d["dt_seq"][n]
        if n == 1:
            p1 = g.get_prims()
            assert np.array_equal(p1, d["p_step1"]), np.abs(p1 - d["p_step1"]).max()
    p = g.get_prims()
    assert np.array_equal(p, d["p_final"]), np.abs(p - d["p_final"]).max()
    if Uo is not None:
        # the carried conserved state, floors' and fixes' corrections included
        assert np.array_equal(g.get_conserved(), Uo)


def test_finite_volume_restart_from_conserved_state():
    """Two steps, then the primitives and the carried conserved state moved
    to a fresh context (as a restart does) and three more steps: the
    reference's final state."""
    d = load("t2d_fv")
    g = make_grid(d)
    g.set_prims(d["p_init"])
    dt = float(d["dt_seq"][0])
    for n in (1, 2):
        dt, _, _ = g.step_ch(dt)
    p, U = g.get_prims(), g.get_conserved()
    h = make_grid(d)
    h.set_prims(p)
    h.set_conserved(U)
    for n in (3, 4, 5):
        dt, _, _ = h.step_ch(dt)
        assert dt == d["dt_seq"][n]
    assert np.array_equal(h.get_prims(), d["p_final"])


@pytest.mark.parametrize("name", ["t2d_weno", "t3d_weno", "t3d_fvweno"])
@pytest.mark.parametrize("shape", [dict(concurrent=0), dict(weno_xmarch=1), dict(weno_xmarch=3, weno_nseg=3, weno_nseg3=2),
                                   dict(weno_xwaves=1, weno_nseg=6, weno_nseg3=4, bs_flux=64, bs_update=128)])
def test_launch_shapes_keep_every_bit(name, shape):
    """Every launch shape of the tuner, the i-direction marching scheme
    included, gives the reference's bits."""
    d = load(name)
    g = make_grid(d)
    t = g.tune()
    for k, v in shape.items():
        setattr(t, k, v)
    g.set_tune(t)
    g.set_prims(d["p_init"])
    dt = float(d["dt_seq"][0])
    for n in range(1, int(d["nsteps"]) + 1):
        dt, nf, it = g.step_ch(dt)
        assert dt == d["dt_seq"][n]
    assert np.array_equal(g.get_prims(), d["p_final"])


SUBSTEP_TABLES = {
    # step_ch.c:442 setup_rktimestep: (CUf, Cunew, pb, pf) per substep
    2: ([(1.0, 0.0, 0.5, 0.0), (1.0, 0.0, 1.0, 0.0)], [(0.0, 0.0, 0.0), (0.0, 0.0, 1.0)], [0, 1], [1, 0]),
    3: ([(1.0, 0.0, 1.0, 0.0), (3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0, 0.0), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.0)],
        [(0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)], [0, 1, 2], [1, 2, 0]),
    4: ([(1.0, 0.0, 0.5, 0.0), (1.0, 0.0, 0.5, 0.0), (1.0, 0.0, 1.0, 0.0), (1.0, 0.0, 1.0, 0.0)],
        [(1.0, 1.0 / 6.0, 0.0), (0.0, 1.0 / 3.0, 0.0), (0.0, 1.0 / 3.0, 0.0), (0.0, 1.0 / 6.0, 0.0)],
        [0, 1, 2, 1], [1, 2, 1, 0]),
}


@pytest.mark.parametrize("name", ["t2d_mc", "t2d_rk3", "t2d_para", "t3d_fv"])
def test_substeps_driven_by_caller_match_reference(name):
    """The reference's step loop (step_ch.c:97-166) with its own RK tables,
    one hg_advance_substep per substep; dt from advance's ndt as step_ch.c
    sets it (min over the substeps' last, SAFE limit)."""
    d = load(name)
    g = make_grid(d)
    g.set_prims(d["p_init"])
    n = int(g.cfg.timeorder) or 2
    cuf, cun, pb, pf = SUBSTEP_TABLES[n]
    dt = float(d["dt_seq"][0])
    for step in range(1, int(d["nsteps"]) + 1):
        for t in range(n):
            ndt = g.advance_substep(0, pb[t], pf[t], cuf[t], cun[t], t, n, dt)
        ndt = min(ndt, 1.3 * dt)  # step_ch.c:243-245, SAFE 1.3
        dt = ndt
        assert dt == d["dt_seq"][step]
    assert np.array_equal(g.get_prims(), d["p_final"])
