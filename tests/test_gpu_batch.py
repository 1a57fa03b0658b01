"""Several tori of different sizes, spins and torus radii advanced by one
HIP graph (hg_graph_create, and hg_graph_create_grouped with one launch per
kernel over all the grids): every grid matches its own single-grid run bit
for bit, on every replay of the graph; and the launch-shape autotuner
(hg_autotune) leaves results unchanged."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [(32, 32, 0.0, 10.0, 1), (48, 64, 0.5, 11.0, 2), (64, 64, 0.9375, 12.0, 3), (32, 48, 0.98, 14.0, 4),
         (64, 32, 0.3, 13.0, 5)]


def torus(n1, n2, a, rmax, seed, scheme="weno"):
    from harm_finalwham_amd import harmgpu
    if scheme == "weno":
        g = harmgpu.Grid(n1, n2, 1, nb=6, lim=harmgpu.HG_LIM_WENO5BND, fluxmethod=harmgpu.HG_FLUX_HLL)
    else:
        # the torus file's own scheme: PARA + HLL + RK4, finite volume
        g = harmgpu.Grid(n1, n2, 1, nb=3, lim=harmgpu.HG_LIM_PARA, fluxmethod=harmgpu.HG_FLUX_HLL, timeorder=4,
                         doenoflux=harmgpu.HG_ENOFINITEVOLUME)
    rh = 1.0 + math.sqrt(1.0 - a * a)
    g.build_mks_geometry(a=a, Rin=0.9 * rh, Rout=40.0, R0=0.0, hslope=0.3)
    g.init_torus(rin=6.0, rmax=rmax, kappa=1e-3, beta=100.0, randfact=4e-2, seed=seed)
    return g


@pytest.mark.parametrize("scheme", ["weno", "fv"])
def test_graph_matches_single_runs_on_every_replay(scheme):
    import torch
    from harm_finalwham_amd import harmgpu
    print("case graph of", len(SPECS), "tori,", scheme, flush=True)
    batch = [torus(*s, scheme=scheme) for s in SPECS]
    alone = [torus(*s, scheme=scheme) for s in SPECS]
    graph = harmgpu.Graph(batch, nsteps=3)
    x = torch.ones(1 << 20, device="cuda", dtype=torch.float64)
    for replay in range(3):
        graph.launch()
        for g in alone:
            g.step_async(3)
        torch.cuda.synchronize()
        for gb, ga in zip(batch, alone):
            assert gb.get_dt() == ga.get_dt(), replay
            assert np.array_equal(gb.get_prims(), ga.get_prims()), replay
            if scheme == "fv":
                assert np.array_equal(gb.get_conserved(), ga.get_conserved()), replay
        # unrelated eager work between replays
        x = x * 1.0000001 + 1.0
        torch.cuda.synchronize()
    graph.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_grouped_graph_matches_single_runs_on_every_replay(mode):
    """hg_graph_create_grouped: one launch per kernel over every grid; each
    grid still evolves bit for bit as alone, with either inversion mode."""
    import torch
    from harm_finalwham_amd import harmgpu
    print("case grouped graph of", len(SPECS), "tori,", mode, flush=True)
    batch = [torus(*s) for s in SPECS]
    alone = [torus(*s) for s in SPECS]
    if mode == "mixed":
        for g in batch + alone:
            g.set_inversion(harmgpu.HG_C2P_MIXED)
    graph = harmgpu.Graph(batch, nsteps=3, grouped=True)
    x = torch.ones(1 << 20, device="cuda", dtype=torch.float64)
    for replay in range(3):
        graph.launch()
        for g in alone:
            g.step_async(3)
        torch.cuda.synchronize()
        for m, (gb, ga) in enumerate(zip(batch, alone)):
            assert gb.get_dt() == ga.get_dt(), (replay, m)
            assert np.array_equal(gb.get_prims(), ga.get_prims()), (replay, m)
            assert np.array_equal(gb.get_pflag(), ga.get_pflag()), (replay, m)
        x = x * 1.0000001 + 1.0  # unrelated eager work between replays
        torch.cuda.synchronize()
    graph.close()


@pytest.mark.parametrize("timeorder", [1, 3, 4])
def test_grouped_graph_other_time_orders_match_single_runs(timeorder):
    """Grouped launches with Euler, TVD RK3 and RK4 substeps (ucum carried
    across them): every grid bit for bit as alone."""
    import torch
    from harm_finalwham_amd import harmgpu

    def make(n1, n2, a, rmax, seed):
        g = harmgpu.Grid(n1, n2, 1, nb=6, lim=harmgpu.HG_LIM_WENO5BND, fluxmethod=harmgpu.HG_FLUX_HLL,
                         timeorder=timeorder)
        rh = 1.0 + math.sqrt(1.0 - a * a)
        g.build_mks_geometry(a=a, Rin=0.9 * rh, Rout=40.0, R0=0.0, hslope=0.3)
        g.init_torus(rin=6.0, rmax=rmax, kappa=1e-3, beta=100.0, randfact=4e-2, seed=seed)
        return g

    print("case grouped graph, timeorder", timeorder, flush=True)
    batch = [make(*sp[:2], *sp[2:]) for sp in SPECS]
    alone = [make(*sp[:2], *sp[2:]) for sp in SPECS]
    graph = harmgpu.Graph(batch, nsteps=2, grouped=True)
    for replay in range(2):
        graph.launch()
        for g in alone:
            g.step_async(2)
        torch.cuda.synchronize()
        for m, (gb, ga) in enumerate(zip(batch, alone)):
            assert gb.get_dt() == ga.get_dt(), (replay, m)
            assert np.array_equal(gb.get_prims(), ga.get_prims()), (replay, m)
    graph.close()


def test_grouped_graph_with_3d_grids_matches_single_runs():
    """Grouped launches over 3-D and 2-D tori together: the k-direction
    classes (reconstruction, flux, periodic ghost fill) run on the 3-D
    members only; every grid bit for bit as alone."""
    import torch
    from harm_finalwham_amd import harmgpu

    def make(n1, n2, n3, a, rmax, seed):
        g = harmgpu.Grid(n1, n2, n3, nb=6, lim=harmgpu.HG_LIM_WENO5BND, fluxmethod=harmgpu.HG_FLUX_HLL)
        rh = 1.0 + math.sqrt(1.0 - a * a)
        g.build_mks_geometry(a=a, Rin=0.9 * rh, Rout=40.0, R0=0.0, hslope=0.3)
        g.init_torus(rin=6.0, rmax=rmax, kappa=1e-3, beta=100.0, randfact=4e-2, seed=seed)
        return g

    specs = [(32, 32, 8, 0.9375, 12.0, 1), (48, 32, 16, 0.5, 11.0, 2), (64, 48, 1, 0.3, 13.0, 3),
             (40, 24, 12, 0.0, 10.5, 4)]
    print("case grouped graph of 2-D and 3-D tori", flush=True)
    batch = [make(*s) for s in specs]
    alone = [make(*s) for s in specs]
    graph = harmgpu.Graph(batch, nsteps=2, grouped=True)
    for replay in range(2):
        graph.launch()
        for g in alone:
            g.step_async(2)
        torch.cuda.synchronize()
        for m, (gb, ga) in enumerate(zip(batch, alone)):
            assert gb.get_dt() == ga.get_dt(), (replay, m)
            assert np.array_equal(gb.get_prims(), ga.get_prims()), (replay, m)
    graph.close()


def test_grouped_graph_rejects_other_schemes():
    from harm_finalwham_amd import harmgpu
    a, b = torus(*SPECS[0]), torus(*SPECS[1])
    fv = torus(*SPECS[2], scheme="fv")
    with pytest.raises(harmgpu.HarmError, match="WENO5BND"):
        harmgpu.Graph([a, fv], nsteps=1, grouped=True)
    b.set_inversion(harmgpu.HG_C2P_MIXED)
    with pytest.raises(harmgpu.HarmError, match="inversion mode"):
        harmgpu.Graph([a, b], nsteps=1, grouped=True)
    with pytest.raises(harmgpu.HarmError):
        harmgpu.Graph([a, a], nsteps=1, grouped=True)


def test_graph_rejects_duplicate_contexts():
    from harm_finalwham_amd import harmgpu
    g = torus(*SPECS[0])
    with pytest.raises(harmgpu.HarmError):
        harmgpu.Graph([g, g], nsteps=1)


@pytest.mark.parametrize("scheme", ["weno", "fv"])
def test_autotune_keeps_results(scheme):
    """The tuner's trials leave no trace: primitives, dt, failure codes and,
    in the finite-volume scheme, the carried conserved state match an
    untuned twin before and after three more steps."""
    from harm_finalwham_amd import harmgpu
    print("case autotune", scheme, flush=True)
    a = torus(*SPECS[2], scheme=scheme)
    b = torus(*SPECS[2], scheme=scheme)
    for g in (a, b):
        g.step_async(2)  # a state that is not the initial one
    t = b.autotune(nsteps=1)
    assert t.bs_flux in (64, 128, 256) and t.weno_xwaves in (1, 2, 4)
    assert b.get_dt() == a.get_dt()
    assert np.array_equal(b.get_prims(), a.get_prims())
    for g in (a, b):
        g.step_async(3)
        assert g.get_dt() > 0
    assert b.get_dt() == a.get_dt()
    assert np.array_equal(b.get_prims(), a.get_prims())
    assert np.array_equal(b.get_pflag(), a.get_pflag())
    if scheme == "fv":
        assert np.array_equal(b.get_conserved(), a.get_conserved())
    bad = harmgpu.HgTune(**{**t.as_dict(), "bs_flux": 100})
    with pytest.raises(harmgpu.HarmError):
        b.set_tune(bad)
