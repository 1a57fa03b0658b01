"""Reading the reference's restart dumps (restart.c:206, text output): the
file restart_write produced after two MC steps of the 32x32 torus is parsed
to exactly the values the reference's own restart_read gets from it."""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(__file__), "golden")


def _file(tmp_path):
    d = np.load(os.path.join(HERE, "rdump_t2d.npz"))
    path = tmp_path / "rdump-0"
    path.write_bytes(d["text"].tobytes())
    return d, str(path)


def test_parse_matches_reference_reader(tmp_path):
    from harm_finalwham_amd import harmgpu
    d, path = _file(tmp_path)
    p, u, info = harmgpu.rdump_parse(path, 32, 32, 1)
    assert np.array_equal(p, d["p_active"])
    assert np.array_equal(u, d["unew"])
    h = d["header"]  # t tf nstep a gam cour dt lim fluxmethod R0 Rin Rout hslope RHOMIN UUMIN GAMMAMAX
    got = [info["t"], info["tf"], info["nstep"], info["a"], info["gam"], info["cour"], info["dt"], info["lim"],
           info["fluxmethod"], info["R0"], info["Rin"], info["Rout"], info["hslope"], info["rhomin"],
           info["uumin"], info["gammamax"]]
    assert np.array_equal(np.array(got, dtype=np.float64), h)
    assert info["n"] == (32, 32, 1)
    # the header as the torus set-up writes it: outflow in r, polar axis, periodic phi
    assert info["timeorder"] == 2 and info["gammafail"] == 2500.0


def test_parse_rejects_other_grids_and_truncation(tmp_path):
    from harm_finalwham_amd import harmgpu
    d, path = _file(tmp_path)
    with pytest.raises(harmgpu.HarmError):
        harmgpu.rdump_parse(path, 64, 32, 1)
    short = tmp_path / "short"
    short.write_bytes(d["text"].tobytes()[: len(d["text"]) // 2])
    with pytest.raises(harmgpu.HarmError):
        harmgpu.rdump_parse(str(short), 32, 32, 1)
    with pytest.raises(harmgpu.HarmError):
        harmgpu.rdump_parse(str(tmp_path / "missing"), 32, 32, 1)


@pytest.mark.gpu
def test_read_into_context(tmp_path):
    """hg_rdump_read: the dump's primitives in the active zones, ghost zones
    from bound_prim (the same as bounding the parsed state on the GPU)."""
    from golden_util import load
    from test_gpu_step import make_grid
    d, path = _file(tmp_path)
    base = load("t2d_mc")
    g = make_grid(base)
    u, info = g.rdump_read(path)
    assert np.array_equal(u, d["unew"]) and info["nstep"] == 2
    p = g.get_prims()
    nb = int(base["params"][3])
    assert np.array_equal(p[nb:nb + 32, nb:nb + 32], d["p_active"])
    ref = make_grid(base)
    full = p.copy()
    full[:nb] = 0.0
    full[nb + 32:] = 0.0
    ref.set_prims(full)
    ref.bound_prim()
    assert np.array_equal(p, ref.get_prims())
