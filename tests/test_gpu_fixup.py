"""post_fixup on the GPU (fixup_checksolution + fixup_utoprim between the two
bound_prim calls) against the reference's own post_fixup, on torus states
with planted inversion failures, fast zones and hot zones, for the first and
the final substep: bit-for-bit primitives and failure codes; and the full
step (which now runs post_fixup) against the oracle on a state with
planted fast zones."""
import numpy as np
import pytest

from test_fixup_golden import expected, load_case
from test_gpu_step import make_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["fixup_t2d.npz", "fixup_t3d.npz"])
@pytest.mark.parametrize("fs", [0, 1])
def test_post_fixup_matches_reference(name, fs):
    print("case post_fixup", name, "finalstep", fs, flush=True)
    d, base, p_in = load_case(name)
    g = make_grid(base)
    g.set_prims(p_in)
    g.post_fixup(codes=d["flags"], finalstep=fs)
    p_ref, pf_ref = expected(d, p_in, fs)
    nb = int(base["params"][3])
    n1, n2, n3 = d["flags"].shape
    b3 = nb if n3 > 1 else 0
    assert np.array_equal(g.get_pflag(), pf_ref[nb:nb + n1, nb:nb + n2, b3:b3 + n3])
    assert np.array_equal(g.get_prims(), p_ref)


def test_post_fixup_without_failures_changes_nothing():
    from golden_util import load
    d = load("t2d_mc")
    g = make_grid(d)
    g.set_prims(d["p_step1"])
    n1, n2, n3 = (int(x) for x in d["params"][:3])
    g.post_fixup(codes=np.zeros((n1, n2, n3), np.int32), finalstep=1)
    assert np.array_equal(g.get_prims(), d["p_step1"])
    assert not g.get_pflag().any()


def test_step_with_fast_zones_matches_oracle():
    """Two RK2 steps from a state with planted fast zones: the neighbour
    check flags some of them in every substep and the average repairs them;
    GPU and oracle agree bit for bit."""
    import pyoracle
    d, base, p_in = load_case("fixup_t2d.npz")
    # keep the fast zones, drop the planted codes (the inversion sets its own)
    g = make_grid(base)
    G = pyoracle.Grid.from_fixture(base)
    p0 = G.bound(p_in)
    g.set_prims(p0)
    dt = float(base["dt_seq"][0]) * 0.25
    po, dto = p0, dt
    for n in range(2):
        dt, nf, it = g.step_ch(dt)
        po, dto, nfo, ito = G.step(po, dto)
        print("step", n, "gpu dt", dt, "oracle dt", dto, "fails", nf, nfo, flush=True)
        assert dt == dto and nf == nfo and it == ito
        assert np.array_equal(g.get_prims(), po)
    # the check did something: some zone ended with the FIXED code
    assert (g.get_pflag() == -1).any()
