"""The library's modified Kerr-Schild geometry builder (host code, no GPU)
against the tables the reference's set_grid (set_grid.c:7) produced for the
golden grids: bit-identical metric, inverse metric, determinant and
connection at every stored location, radii and spacings."""
import numpy as np
import pytest

from golden_util import load


@pytest.mark.parametrize("name", ["t2d_mc", "t3d_weno"])
def test_mks_geometry_tables_match_reference(name):
    from harm_finalwham_amd import harmgpu
    d = load(name)
    par = d["params"]
    n1, n2, n3, nb = (int(x) for x in par[:4])
    g = harmgpu.mks_geometry_tables(n1, n2, n3, nb, a=par[4], Rin=par[6], Rout=par[7], R0=par[9],
                                    hslope=par[8])
    for key in ("gcov", "gcon", "gdet", "conn"):
        assert g[key].shape == d[key].shape, key
        assert np.array_equal(g[key], d[key]), key
    assert np.array_equal(g["gcovpert"][..., 0], d["gcovpert"][..., 0])
    assert np.array_equal(g["r_cent"], d["xv_cent"][:, :, 1, 1])
    assert np.array_equal(g["dx"][1:], par[12:15])
    assert g["dx"][1] == np.log((par[7] - par[9]) / (par[6] - par[9])) / n1


def test_mks_geometry_invariants():
    """Properties the geometry must have whatever the reference: g^{-1} g = 1,
    gdet^2 = -det g, symmetric connection in its lower indices."""
    from harm_finalwham_amd import harmgpu
    g = harmgpu.mks_geometry_tables(16, 24, 1, 3)
    prod = np.einsum("ijlab,ijlbc->ijlac", g["gcon"], g["gcov"])
    assert np.max(np.abs(prod - np.eye(4))) < 1e-9
    det = np.linalg.det(g["gcov"])
    assert np.max(np.abs(g["gdet"] ** 2 / -det - 1)) < 1e-12
    c = g["conn"]
    # centred differences of g carry ~1e-11 relative rounding
    assert np.max(np.abs(c - np.swapaxes(c, 3, 4))) <= 1e-10 * np.max(np.abs(c))


def test_mks_geometry_rejects_bad_arguments():
    from harm_finalwham_amd import harmgpu
    with pytest.raises(harmgpu.HarmError):
        harmgpu.mks_geometry_tables(16, 16, 1, 3, Rin=0.0, R0=0.0)


@pytest.mark.parametrize("name", ["c1_mc_100", "f3d_weno_1"])
def test_mks_geometry_digests_match_reference(name):
    """The 128x128 and 256x128 (x64) grids: digests of the reference's tables.
    These grids reach angles where a libm sincos() call and separate sin()
    and cos() calls round differently in the last bit."""
    import os
    from golden_util import GOLDEN, state_digest
    from harm_finalwham_amd import harmgpu
    if not os.path.exists(os.path.join(GOLDEN, name + ".npz")):
        pytest.skip("fixture not generated")
    d = load(name)
    par = d["params"]
    n1, n2, n3, nb = (int(x) for x in par[:4])
    g = harmgpu.mks_geometry_tables(n1, n2, n3, nb, a=par[4], Rin=par[6], Rout=par[7], R0=par[9],
                                    hslope=par[8])
    for key in ("gcov", "gcon", "gdet", "conn"):
        assert state_digest(g[key]) == str(d["digest_" + key]), key
