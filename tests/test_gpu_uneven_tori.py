"""48 independent 2-D tori of uneven sizes advanced by one HIP graph
(hg_graph_create: a branch per grid; hg_graph_create_grouped: one launch
per kernel over all the grids), at the batch's own sizes: {64x64, 96x128, 128x128,
192x192, 256x256}, spins a ~ U[0, 0.98], pressure-maximum radii
r_pmax ~ U[10, 14], seeded perturbations, WENO5BND + HLL + flux-CT + Noble 2D
with RK2 (each instance is one grid's step_ch, step_ch.c:35).

Every grid of the graph matches its own single-grid run bit for bit after
each of two replays (eager work in between), and one 128x128 member (a =
0.53, r_pmax = 12, R_in = 0.9 r_+) matches the reference build's own run of
the same torus over the same ten steps (fixture c5_t2d_128: every dt, the
final state, constotal and divbmaxavg)."""
import math

import numpy as np
import pytest

from golden_util import load, state_digest

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (96, 128), (128, 128), (192, 192), (256, 256)]
PINNED = 2  # index of the member pinned to the reference fixture


def batch_specs(n=48, seed=5):
    """(n1, n2), a, r_pmax, perturbation seed; the first five members take
    each size once, the pinned one the reference fixture's parameters."""
    rng = np.random.default_rng(seed)
    specs = []
    for m in range(n):
        size = SIZES[m] if m < len(SIZES) else SIZES[int(rng.integers(len(SIZES)))]
        specs.append((size, float(rng.uniform(0.0, 0.98)), float(rng.uniform(10.0, 14.0)),
                      int(rng.integers(1, 1 << 30))))
    d = load("c5_t2d_128")
    specs[PINNED] = ((128, 128), float(d["a"]), 12.0, 7)
    return specs


def make(spec):
    from harm_finalwham_amd import harmgpu
    (n1, n2), a, rmax, sd = spec
    g = harmgpu.Grid(n1, n2, 1, nb=6, lim=harmgpu.HG_LIM_WENO5BND, fluxmethod=harmgpu.HG_FLUX_HLL)
    rh = 1.0 + math.sqrt(1.0 - a * a)
    g.build_mks_geometry(a=a, Rin=0.9 * rh, Rout=40.0, R0=0.0, hslope=0.3)
    dt = g.init_torus(rin=6.0, rmax=rmax, kappa=1e-3, beta=100.0, randfact=4e-2, seed=sd)
    return g, dt


@pytest.mark.parametrize("grouped", [False, True])
def test_48_uneven_tori_graph_matches_single_runs_and_reference(grouped):
    import torch
    from harm_finalwham_amd import harmgpu
    specs = batch_specs()
    print("case 48 uneven tori, grouped" if grouped else "case 48 uneven tori, one branch per grid",
          {f"{a}x{b}": sum(1 for s in specs if s[0] == (a, b)) for a, b in SIZES}, flush=True)
    d = load("c5_t2d_128")
    made = [make(s) for s in specs]
    batch = [g for g, _ in made]
    alone = [make(s)[0] for s in specs]
    gp, dt0 = made[PINNED]
    assert abs(0.9 * (1.0 + math.sqrt(1.0 - d["a"] ** 2)) - d["Rin"]) < 1e-15
    assert dt0 == d["dt_seq"][0]
    assert state_digest(gp.get_prims()) == str(d["digest_init"])
    nsteps = 5
    graph = harmgpu.Graph(batch, nsteps=nsteps, grouped=grouped)
    x = torch.ones(1 << 20, device="cuda", dtype=torch.float64)
    for replay in range(2):
        graph.launch()
        for g in alone:
            g.step_async(nsteps)
        torch.cuda.synchronize()
        for m, (gb, ga) in enumerate(zip(batch, alone)):
            assert gb.get_dt() == ga.get_dt(), (replay, m)
            assert np.array_equal(gb.get_prims(), ga.get_prims()), (replay, m)
        x = x * 1.0000001 + 1.0  # unrelated eager work between replays
        torch.cuda.synchronize()
        print("replay", replay, "bit-identical to the single-grid runs", flush=True)
    graph.close()
    # the pinned member against the reference build over the same ten steps
    assert gp.get_dt() == d["dt_seq"][2 * nsteps]
    assert not d["fail_seq"].any()
    p = gp.get_prims()
    assert np.array_equal(p, d["p_final"]), float(np.abs(p - d["p_final"]).max())
    diag = gp.diagnostics(0.0)
    assert np.array_equal(diag["cons"], d["cons"])
    # divbmaxavg normalises with an x87 long double THIRD in the reference
    assert abs(diag["divbmax"] - d["divb"][0]) <= 1e-14 * abs(d["divb"][0])
    assert abs(diag["divbavg"] - d["divb"][1]) <= 1e-13 * abs(d["divb"][1])
    for g in batch + alone:
        g.close()
