"""The reference's restart dumps (restart.c:206, text output): the files
restart_write produced after two steps of the 32x32 torus (MC + RK2, and the
torus file's finite-volume PARA + RK4 scheme) are parsed to exactly the
values the reference's own restart_read gets from them, and the library
writes the same bytes after the same steps."""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(__file__), "golden")


def _file(tmp_path, name="rdump_t2d"):
    d = np.load(os.path.join(HERE, name + ".npz"))
    path = tmp_path / "rdump-0"
    path.write_bytes(d["text"].tobytes())
    return d, str(path)


@pytest.mark.parametrize("name", ["rdump_t2d", "rdump_t2d_fv"])
def test_parse_matches_reference_reader(tmp_path, name):
    from harm_finalwham_amd import harmgpu
    d, path = _file(tmp_path, name)
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
    assert info["timeorder"] == (4 if name.endswith("_fv") else 2) and info["gammafail"] == 2500.0


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


@pytest.mark.gpu
@pytest.mark.parametrize("name,base_name", [("rdump_t2d", "t2d_mc"), ("rdump_t2d_fv", "t2d_fv")])
def test_write_matches_reference_bytes(tmp_path, name, base_name):
    """hg_rdump_write after the same two steps the reference took before its
    restart_write: the same bytes (the header line passed through).  The
    finite-volume dump's unew is the carried conserved state."""
    from golden_util import load
    from test_gpu_step import make_grid
    print("case rdump write", name, flush=True)
    d = load(name)
    text = bytes(d["text"])
    header = text.split(b"\n", 1)[0]
    base = load(base_name)
    g = make_grid(base)
    g.set_prims(base["p_init"])
    dt = float(base["dt_seq"][0])
    for n in (1, 2):
        dt, _, _ = g.step_ch(dt)
    out = tmp_path / "rdump-0"
    g.rdump_write(str(out), header)
    assert out.read_bytes() == text


@pytest.mark.gpu
def test_finite_volume_read_takes_unew(tmp_path):
    from golden_util import load
    from test_gpu_step import make_grid
    d, path = _file(tmp_path, "rdump_t2d_fv")
    base = load("t2d_fv")
    g = make_grid(base)
    u, info = g.rdump_read(path)
    assert np.array_equal(u, d["unew"]) and info["timeorder"] == 4
    nb = int(base["params"][3])
    U = g.get_conserved()
    assert np.array_equal(U[nb:nb + 32, nb:nb + 32], d["unew"])


@pytest.mark.gpu
def test_write_needs_a_step(tmp_path):
    from harm_finalwham_amd import harmgpu
    from golden_util import load
    from test_gpu_step import make_grid
    base = load("t2d_mc")
    g = make_grid(base)
    g.set_prims(base["p_init"])
    with pytest.raises(harmgpu.HarmError):
        g.rdump_write(str(tmp_path / "x"), "header")
