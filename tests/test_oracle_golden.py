"""The CPU oracle against the golden vectors of the reference build
(tests/golden/make_golden.py): bit-for-bit agreement."""
import numpy as np
import pytest

import pyoracle as po
from golden_util import active, load


@pytest.mark.parametrize("name", ["c2p_g43", "c2p_g53"])
def test_c2p_matches_reference(name):
    d = load(name)
    n = d["U"].shape[0]
    U = po.primtoU_batch(d["gcov"].reshape(n, 16), d["gcon"].reshape(n, 16), d["gdet"], d["gcovpert"],
                         d["p_true"], float(d["gam"]))
    assert np.array_equal(U, d["U"])  # phys.c:225 primtoU
    p, fl, it = po.c2p_batch(d["gcov"].reshape(n, 16), d["gcon"].reshape(n, 16), d["gdet"], d["gcovpert"],
                             d["U"], float(d["gam"]), d["p_guess"])
    assert np.array_equal(p, d["p_out"])
    assert np.array_equal(fl, d["pflag"])
    assert np.array_equal(it, d["iters"])


def test_c2p_failure_codes_present():
    # the fixtures exercise at least one failure branch besides success
    d = load("c2p_g43")
    assert set(np.unique(d["pflag"])) - {0}


TORI = ["t2d_mc", "t2d_weno", "t3d_weno", "t2d_para", "t3d_para", "t2d_fv", "t3d_fv", "t2d_rk3", "t2d_euler",
        "t2d_force", "t2d_musta", "t2d_hlllaxf1", "t2d_fvweno", "t3d_fvweno", "t2d_fv_30", "t3d_weno_10", "t3d_fv_10",
        "t2d_vanl", "t2d_minm", "t2d_donor"]


@pytest.mark.parametrize("name", TORI)
def test_torus_steps_match_reference(name):
    d = load(name)
    G = po.Grid.from_fixture(d)
    nb, n3 = int(d["params"][3]), int(d["params"][2])
    p, dt = d["p_init"], float(d["dt_seq"][0])
    U = G.cons_from_prims(p) if G.g.fv else None
    for n in range(1, int(d["nsteps"]) + 1):
        if U is None:
            p, dt, nf, _ = G.step(p, dt)
        else:
            p, U, dt, nf, _ = G.step_fv(p, U, dt)
        # a step whose final flags hold failures failed in some substep
        if d["fail_seq"][n]:
            assert nf > 0
        assert dt == d["dt_seq"][n]
        if n == 1:
            assert np.array_equal(p, d["p_step1"])
    assert np.array_equal(p, d["p_final"])


@pytest.mark.parametrize("name", TORI)
def test_fluxcalc_matches_reference(name):
    d = load(name)
    G = po.Grid.from_fixture(d)
    nb = int(d["params"][3])
    F, ndt = G.fluxcalc(d["p_init"])
    assert np.array_equal(ndt, d["ndt_init"])
    # faces read by the active-zone update: i=0..n1 (dir 1), j=0..n2 (dir 2), k=0..n3 (dir 3)
    n3 = int(d["params"][2])
    k = slice(nb, -nb) if n3 > 1 else slice(None)
    kf = slice(nb, -nb + 1) if n3 > 1 else slice(None)
    assert np.array_equal(F[0][nb:-nb + 1, nb:-nb, k], d["F1_init"][nb:-nb + 1, nb:-nb, k])
    assert np.array_equal(F[1][nb:-nb, nb:-nb + 1, k], d["F2_init"][nb:-nb, nb:-nb + 1, k])
    if n3 > 1:
        assert np.array_equal(F[2][nb:-nb, nb:-nb, kf], d["F3_init"][nb:-nb, nb:-nb, kf])


def test_ghost_fill_is_identity_on_reference_state():
    # the reference state after a step is already bounded; bound_prim must reproduce it
    d = load("t2d_mc")
    G = po.Grid.from_fixture(d)
    assert np.array_equal(G.bound(d["p_step1"]), d["p_step1"])


@pytest.mark.parametrize("name", ["t2d_weno", "t3d_weno", "t2d_para", "t3d_para", "t3d_fv"])
def test_interface_states_match_reference(name):
    """WENO5BND (slope_lim_linetype_c2e, interpline.c:517) and PARA (para4,
    interppoint.c:738) states of cells -1..n along each direction."""
    d = load(name)
    G = po.Grid.from_fixture(d)
    nb = int(d["params"][3])
    for dr in (1, 2, 3):
        if "pleft_dir%d" % dr not in d:
            continue
        pl, pr = G.reconstruct(d["p_init"], dr)
        sl = [slice(None)] * 3
        sl[dr - 1] = slice(nb - 1, -nb + 1)
        sl = tuple(sl)
        assert np.array_equal(pl[sl], d["pleft_dir%d" % dr][sl])
        assert np.array_equal(pr[sl], d["pright_dir%d" % dr][sl])


def test_finite_volume_step_carries_the_conserved_state():
    """Planted mistake: deriving Ui from the primitives each step (the
    standard scheme) instead of carrying ucum departs from the reference's
    finite-volume run within its five steps."""
    d = load("t2d_fv")
    G = po.Grid.from_fixture(d)
    G.g.fv = 0
    p, dt = d["p_init"], float(d["dt_seq"][0])
    for n in range(1, int(d["nsteps"]) + 1):
        p, dt, _, _ = G.step(p, dt)
    assert not np.array_equal(p, d["p_final"])
