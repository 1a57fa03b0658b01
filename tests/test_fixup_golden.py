"""post_fixup (fixup.c:37: bound_pflag, fixup_checksolution, fixup_utoprim)
of the oracle against the reference's own post_fixup, on torus states with
planted inversion failures, fast zones (gamma 3-8) and hot zones (u x 30),
for the first and the final substep (tests/golden/make_golden.py --fixup)."""
import os

import numpy as np
import pytest

import pyoracle

HERE = os.path.join(os.path.dirname(__file__), "golden")


def load_case(name):
    d = np.load(os.path.join(HERE, name))
    base = np.load(os.path.join(HERE, str(d["base"])))
    p_in = base["p_init"].copy().reshape(-1)
    p_in[d["mod_idx"]] = d["mod_val"]
    return d, base, p_in.reshape(base["p_init"].shape)


def expected(d, p_in, fs):
    out = p_in.copy().reshape(-1)
    out[d["fs%d_idx" % fs]] = d["fs%d_val" % fs]
    return out.reshape(p_in.shape), d["fs%d_pflag" % fs].astype(np.int32)


@pytest.mark.parametrize("name", ["fixup_t2d.npz", "fixup_t3d.npz"])
@pytest.mark.parametrize("fs", [0, 1])
def test_oracle_post_fixup_matches_reference(name, fs):
    d, base, p_in = load_case(name)
    G = pyoracle.Grid.from_fixture(base)
    p, pf = G.post_fixup_sequence(p_in, d["flags"], fs)
    p_ref, pf_ref = expected(d, p_in, fs)
    assert np.array_equal(pf, pf_ref)
    assert np.array_equal(p, p_ref)


@pytest.mark.parametrize("name", ["fixup_t2d.npz", "fixup_t3d.npz"])
def test_planted_cases_reach_every_branch(name):
    """The fixtures exercise what they are meant to: averaged failures,
    fast zones flagged by the neighbour check, negative-density codes left
    alone before the final substep and counted on it."""
    d, base, p_in = load_case(name)
    flags = d["flags"]
    nb = int(base["params"][3])
    n3 = flags.shape[2]
    b3 = nb if n3 > 1 else 0
    for fs in (0, 1):
        p_ref, pf_ref = expected(d, p_in, fs)
        act = pf_ref[nb:nb + flags.shape[0], nb:nb + flags.shape[1], b3:b3 + n3]
        # converged-failure codes become FIXED (-1) on both substeps
        assert np.all(act[flags == 101] == -1)
        # codes 9/10/11 are kept before the final substep and reset on it
        neg = np.isin(flags, (9, 10, 11))
        if fs == 0:
            assert np.array_equal(act[neg], flags[neg])
        else:
            assert np.all(act[neg] == -1)
        # some zones that inverted fine were flagged by the check and fixed
        fixed_good = (flags == 0) & (act == -1)
        assert fixed_good.sum() > 0
        changed = np.zeros(p_in.shape[:3], bool)
        changed.reshape(-1)[np.unique(d["fs%d_idx" % fs] // 8)] = True
        assert np.all(changed[nb:nb + flags.shape[0], nb:nb + flags.shape[1], b3:b3 + n3][fixed_good])


def test_planted_mistake_is_caught():
    """A post_fixup that forgets the neighbour check fails the fixture."""
    d, base, p_in = load_case("fixup_t2d.npz")
    G = pyoracle.Grid.from_fixture(base)
    # dropping the checked zones: treat the inversion codes as final (no fast
    # zone is flagged) by running on a state whose fast zones are slowed down
    slow = p_in.copy()
    p_ref, _ = expected(d, p_in, 1)
    nb = int(base["params"][3])
    fast = np.zeros(slow.shape[:3], bool)
    fast[nb:-nb, nb:-nb, :] = True
    v = slow[..., 2]
    v[fast & (np.abs(v) > 1.0)] *= 1e-3
    p, _ = G.post_fixup_sequence(slow, d["flags"], 1)
    assert not np.array_equal(p, p_ref)


def _gamma(p, gcov_cent):
    u = p[..., 2:5]
    g = gcov_cent[..., 1:4, 1:4]
    return np.sqrt(1.0 + np.einsum("...a,...ab,...b->...", u, g, u))


@pytest.mark.parametrize("name", ["fixup_t2d.npz", "fixup_t3d.npz"])
def test_order_dependent_pair(name):
    """The check runs in the reference's loop order and sees flags it has
    already set: zone B = (i, j+1) keeps its solution because its earlier
    neighbour A = (i, j) was flagged first; counted with A as a good
    neighbour, B would have 5 of 8 votes and be flagged too."""
    d, base, p_in = load_case(name)
    G = pyoracle.Grid.from_fixture(base)
    nb = int(base["params"][3])
    n1, n2, n3 = d["flags"].shape
    b3 = nb if n3 > 1 else 0
    pb = G.bound(p_in)
    gam = _gamma(pb[:, :, :, :], base["gcov"][:, :-1, 6][:pb.shape[0], :pb.shape[1], None])
    ic, jc = nb + int(0.5 * n1), nb + int(0.5 * n2)
    for fs in (0, 1):
        _, pf_ref = expected(d, p_in, fs)
        for k in ([0, n3 - 1, n3 // 2] if n3 > 1 else [0]):
            kk = k + b3
            assert pf_ref[ic, jc, kk] == -1          # A flagged by the check, then fixed
            assert pf_ref[ic, jc + 1, kk] == 0       # B kept
            i, j, gb = ic, jc + 1, gam[ic, jc + 1, kk]
            nbrs = [(i, j + 1), (i, j - 1), (i + 1, j), (i - 1, j), (i + 1, j + 1), (i + 1, j - 1), (i + 1, j + 1),
                    (i - 1, j - 1)]
            r = np.array([gam[a, b, kk] / gb for a, b in nbrs])
            votes = np.sum((r > 2.0) | (r < 0.5))
            assert votes == 5
