#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference build.

The reference (pseudotensor/harm_finalwham) is compiled from its own C
sources by oracle/refbuild/build_ref.sh into oracle/_ref/harness_<tag>; this
script runs those executables and stores their inputs and outputs as .npz
files.  It needs /root/reference (only in the build container); the GPU box
uses the committed .npz files.

Fixtures
  c2p_g43.npz / c2p_g53.npz   batched Noble-2D inversions (Gamma 4/3, 5/3):
        metric, true primitives, guess -> U (reference primtoU), inverted
        primitives, failure code, Newton iterations (utoprimgen.c:23)
  t2d_mc.npz     2-D torus 32x32, MC + LaxF + flux-CT, 5 RK2 steps
  t2d_weno.npz   2-D torus 32x32, WENO5BND + HLL + flux-CT, 5 RK2 steps
  t3d_weno.npz   3-D torus 16x16x8, WENO5BND + HLL + flux-CT, 1 RK2 step
  t2d_para.npz   2-D torus 32x32, PARA + HLL + flux-CT, 5 RK4 steps
  t3d_para.npz   3-D torus 16x16x8, PARA + HLL + flux-CT, 1 RK4 step
  t2d_fv.npz     2-D torus 32x32, the same with the finite-volume step
        (ENOFINITEVOLUME, avgscheme DONOR), 5 RK4 steps
  t3d_fv.npz     3-D torus 16x16x8, finite-volume PARA + HLL, 2 RK4 steps
  t2d_rk3.npz    2-D torus 32x32, MC + LaxF, 5 TVD RK3 steps
  t2d_euler.npz  2-D torus 32x32, MC + LaxF, 5 Euler steps
  t2d_force.npz  2-D torus 32x32, MC + FORCE, 5 RK2 steps
  t2d_musta.npz  2-D torus 32x32, PARA + MUSTA, 5 RK2 steps
  t2d_hlllaxf1.npz 2-D torus 32x32, WENO5BND + HLLLAXF1, 5 RK2 steps
  t2d_fvweno.npz 2-D torus 32x32, WENO5BND + HLL + RK2, finite volume, 3 steps
  t3d_fvweno.npz 3-D torus 16x16x8, the same, 1 step
  t2d_fv_30.npz  2-D torus 32x32, the torus file's own scheme, 30 steps
  t3d_weno_10.npz 3-D torus 16x16x8, WENO5BND + HLL + RK2, 10 steps
  t3d_fv_10.npz  3-D torus 16x16x8, the torus file's own scheme, 10 steps
  t2d_vanl.npz / t2d_minm.npz / t2d_donor.npz  2-D torus 32x32, VANL + LaxF,
        MINM + HLL, DONOR + LaxF, 5 RK2 steps
        (these seventeen written alone by --para)
Each torus fixture carries the reference geometry tables, the initial
primitives, the reconstruction states (or MC slopes) per direction, the
fluxes after one fluxcalc()+flux_ct() and the primitives after steps.

With --long (and --full):
  c1_mc_100.npz    2-D torus 128x128, MC + LaxF, 100 RK2 steps: dt and
        failure sequences, digests of the initial and final states and of
        the geometry tables, the final state
  f3d_weno_1.npz   3-D torus 256x128x64, WENO5BND + HLL, 1 RK2 step: dt,
        digests (states, geometry) and per-primitive sums only (the states
        are 182 MB)

With --long50:
  f3d_weno_50.npz  3-D torus 256x128x64, WENO5BND + HLL, 50 RK2 steps (about
        an hour of one core here): dt and failure sequences, digests of the
        initial and final states, per-primitive sums of the final state, and
        the reference's constotal (dump_ener.c:596) and divbmaxavg
        (dump_ener.c:482) of the final state

With --c5:
  c5_t2d_128.npz   2-D torus 128x128, WENO5BND + HLL, 10 RK2 steps, at the spin
        and inner radius of one member of the 48-torus batch (a = 0.53,
        Rin = 0.9 of the horizon radius; r_pmax 12 is the torus file's own):
        dt and failure sequences, the final state
"""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from c2p_sample import ks_metric, sample_c2p  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
BUILD = os.path.join(REPO, "oracle", "refbuild", "build_ref.sh")
REFDIR = os.path.join(REPO, "oracle", "_ref")

C2P_SEED = 20240611
LOCS = [0, 1, 2, 3, 4, 5, 7]  # FACE1..3, CORN1..3, CENT (global.h:946-953); CORNT unused


def edge_cases(rng):
    """Hand-made records that exercise the failure branches of
    Utoprim_new_body (utoprim_2d_final.c:218-431)."""
    gcov, gcon, gdet, gpert, p, guess = sample_c2p(8, rng)
    guess[0, 2:5] *= 1e4      # utsq of the guess beyond UTSQ_TOO_BIG -> code 2
    p[1, 1] = 0.0             # cold gas
    guess[2] = p[2]           # exact guess
    p[3, 5:8] = 0.0           # unmagnetised
    guess[3] = p[3]
    p[4, 2:5] = 0.0           # fluid at rest in the normal frame's coordinate basis
    p[5, 1] = 1e-12           # near-cold, highly magnetised
    return gcov, gcon, gdet, gpert, p, guess


def run_c2p(harness, gam, n, seed):
    rng = np.random.default_rng(seed)
    parts = [sample_c2p(n - 8, rng), edge_cases(rng)]
    gcov, gcon, gdet, gpert, p, guess = [np.concatenate(x) for x in zip(*parts)]
    m = p.shape[0]
    rec = np.concatenate([gcov.reshape(m, 16), gcon.reshape(m, 16), gdet[:, None], gpert, p, guess],
                         axis=1)
    assert rec.shape[1] == 53
    with tempfile.TemporaryDirectory() as tmp:
        fin = os.path.join(tmp, "in.bin")
        fout = os.path.join(tmp, "out.bin")
        rec.astype("<f8").tofile(fin)
        subprocess.run([harness, "c2p", fin, fout, repr(gam)], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        out = np.fromfile(fout, dtype="<f8").reshape(m, 18)
    return dict(gcov=gcov, gcon=gcon, gdet=gdet, gcovpert=gpert, p_true=p, p_guess=guess,
                U=out[:, :8], p_out=out[:, 8:16], pflag=out[:, 16].astype(np.int32),
                iters=out[:, 17].astype(np.int32), gam=np.float64(gam), seed=np.int64(seed))


def run_torus(harness, weno, nsteps):
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "out")
        subprocess.run([harness, "torus", out, str(weno), str(nsteps)], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        d = {f[:-4]: np.load(os.path.join(out, f)) for f in os.listdir(out) if f.endswith(".npy")}
    # keep the geometry locations the hot path reads
    for k in ("gcov", "gcon", "gcovpert", "gdet"):
        d[k] = np.ascontiguousarray(d[k][:, :, LOCS])
    d["locs"] = np.array(LOCS, dtype=np.int32)
    d["weno"] = np.int32(weno)
    d["nsteps"] = np.int32(nsteps)
    return d


def state_digest(p):
    """SHA-256 of a state array with -0.0 folded into +0.0 (array_equal
    semantics), for comparing states too large to keep as fixtures."""
    a = np.ascontiguousarray(p, dtype="<f8") + 0.0
    return hashlib.sha256(a.tobytes()).hexdigest()


def run_torus_long(harness, weno, nsteps, keep_final):
    """A longer or larger run: the parameters, the dt and failure sequences,
    digests of the initial and final states, per-primitive sums of the final
    state, and the final state itself when keep_final."""
    d = run_torus(harness, weno, nsteps)
    out = {k: d[k] for k in ("params", "dt_seq", "fail_seq", "weno", "nsteps")}
    out["digest_init"] = np.array(state_digest(d["p_init"]))
    out["digest_final"] = np.array(state_digest(d["p_final"]))
    out["sum_final"] = d["p_final"].reshape(-1, 8).sum(axis=0)
    for k in ("gcov", "gcon", "gdet", "conn"):
        out["digest_" + k] = np.array(state_digest(d[k]))
    if keep_final:
        out["p_final"] = d["p_final"]
    return out


def main_long(full):
    """128x128 MC/LaxF torus over 100 steps; with full, also the 256x128x64
    WENO5BND/HLL torus over one step (about ten minutes and 14 GB here);
    with --fv alone, the 256x128x64 torus in the torus file's own scheme
    (PARA + HLL + RK4, finite volume) over one step."""
    runs = [("c1", (128, 128, 1), 0, 100, True, "c1_mc_100.npz")]
    if full:
        runs.append(("f3d", (256, 128, 64), 1, 1, False, "f3d_weno_1.npz"))
    if "--fv" in sys.argv:
        runs = [("f3d", (256, 128, 64), 3, 1, False, "f3d_fv_1.npz")]
    for tag, (n1, n2, n3), weno, nsteps, keep, name in runs:
        subprocess.run([BUILD, tag, str(n1), str(n2), str(n3)], check=True, stdout=subprocess.DEVNULL)
        d = run_torus_long(os.path.join(REFDIR, "harness_" + tag), weno, nsteps, keep)
        np.savez_compressed(os.path.join(HERE, name), **d)
        print(name, os.path.getsize(os.path.join(HERE, name)))


def run_long(harness, scheme, nsteps, env=None):
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "out")
        subprocess.run([harness, "long", out, str(scheme), str(nsteps)], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       env=dict(os.environ, **(env or {})))
        d = {f[:-4]: np.load(os.path.join(out, f)) for f in os.listdir(out) if f.endswith(".npy")}
    d["scheme"] = np.int32(scheme)
    d["nsteps"] = np.int32(nsteps)
    return d


def main_long50():
    """The benchmark grid at the configuration's stated length: 50 RK2 steps
    of the 256x128x64 WENO5BND/HLL torus."""
    subprocess.run([BUILD, "f3d", "256", "128", "64"], check=True, stdout=subprocess.DEVNULL)
    d = run_long(os.path.join(REFDIR, "harness_f3d"), 1, 50)
    out = {k: d[k] for k in ("dt_seq", "fail_seq", "cons", "divb", "scheme", "nsteps")}
    out["digest_init"] = np.array(state_digest(d["p_init"]))
    out["digest_final"] = np.array(state_digest(d["p_final"]))
    out["sum_final"] = d["p_final"].reshape(-1, 8).sum(axis=0)
    np.savez_compressed(os.path.join(HERE, "f3d_weno_50.npz"), **out)


C5_A = 0.53


def c5_rin(a):
    """0.9 of the outer horizon radius r_+ = 1 + sqrt(1 - a^2)."""
    return 0.9 * (1.0 + np.sqrt(1.0 - a * a))


def main_c5():
    """One 128x128 member of the uneven 48-torus batch, WENO5BND + HLL."""
    subprocess.run([BUILD, "c5", "128", "128", "1"], check=True, stdout=subprocess.DEVNULL)
    env = {"HR_A": repr(C5_A), "HR_RIN": repr(float(c5_rin(C5_A)))}
    d = run_long(os.path.join(REFDIR, "harness_c5"), 1, 10, env)
    out = {k: d[k] for k in ("dt_seq", "fail_seq", "cons", "divb", "scheme", "nsteps", "p_final")}
    out["a"] = np.float64(C5_A)
    out["Rin"] = np.float64(c5_rin(C5_A))
    out["digest_init"] = np.array(state_digest(d["p_init"]))
    np.savez_compressed(os.path.join(HERE, "c5_t2d_128.npz"), **out)


FIXUP_CODES = np.array([101, 101, 101, 2, 3, 4, 9, 10, 11, 14, 15], np.int32)


def plant_fixup(base, rng):
    """A torus state with planted inversion failures and fast, hot zones for
    post_fixup: returns (flat indices, new values) of the primitive entries
    changed and the inversion codes of the active zones.  Planted: random
    failures (4%) with every code the inversion or a fixup can leave, failed
    clusters (2x2, and 3x3 whose centre has no good neighbour), failures on
    the boundary rows, zones with gamma in [3, 8] (alone, in rows along j,
    in 2x2 patches, next to failures), and hot zones (u x 30) among them."""
    m3 = base.shape[2]
    g = np.load(os.path.join(HERE, "t2d_mc.npz" if m3 == 1 else "t3d_weno.npz"))
    n1 = int(g["params"][0]); n2 = int(g["params"][1]); n3 = int(g["params"][2]); nb = int(g["params"][3])
    b3 = nb if n3 > 1 else 0
    gcov11 = g["gcov"][:, :, 6, 1, 1]  # CENT
    p = base.copy()
    flags = np.zeros((n1, n2, n3), np.int32)
    nz = n1 * n2 * n3
    pick = rng.choice(nz, size=max(4, nz // 25), replace=False)
    flags.flat[pick] = rng.choice(FIXUP_CODES, size=pick.size)
    ks = [0, n3 - 1, n3 // 2] if n3 > 1 else [0]
    q = lambda f, n: int(f * n)  # planted positions scale with the grid
    for k in ks:
        flags[q(.12, n1):q(.12, n1) + 2, q(.31, n2):q(.31, n2) + 2, k] = 101       # 2x2 cluster
        flags[q(.31, n1):q(.31, n1) + 3, q(.62, n2):q(.62, n2) + 3, k] = 101       # 3x3 cluster
        flags[0, q(.1, n2), k] = 101                                               # boundary rows
        flags[n1 - 1, q(.16, n2), k] = 2
        flags[q(.22, n1), 0, k] = 101
        flags[q(.28, n1), n2 - 1, k] = 15
    fast = []
    for k in ks:
        a, b = q(.62, n1), q(.19, n2)
        fast += [(a, b, k), (a + 1, b, k), (a, b + 1, k), (a + 1, b + 1, k)]        # 2x2 patch
        fast += [(q(.78, n1), j, k) for j in range(q(.37, n2), q(.37, n2) + 4)]     # row along j
        fast += [(q(.47, n1), q(.81, n2), k), (q(.09, n1), q(.25, n2), k),          # alone, at the
                 (n1 - 1, n2 - 1, k), (q(.12, n1) + 1, q(.31, n2) + 1, k)]          # edge, in a cluster
    gam_t = rng.uniform(3.0, 8.0, size=len(fast))
    for (i, j, k), G in zip(fast, gam_t):
        ii, jj, kk = i + nb, j + nb, k + b3
        p[ii, jj, kk, 2] = np.sqrt(G * G - 1.0) / np.sqrt(gcov11[ii, jj])
        p[ii, jj, kk, 3] = 0.0
        p[ii, jj, kk, 4] = 0.0
        if (i, j) != (q(.12, n1) + 1, q(.31, n2) + 1):
            flags[i, j, k] = 0
    for (i, j, k) in fast[::3]:
        p[i + nb, j + nb, k + b3, 1] *= 30.0
    i0, j0 = q(.47, n1), q(.81, n2)
    flags[i0 - 1, j0, ks[0]] = 101                           # a failed (i-1) neighbour of a fast zone
    flags[i0, j0 + 1, ks[-1]] = 101                          # a failed (i, j+1) neighbour: no votes
    # an order-dependent pair: A = (i, j) with gamma 8 is flagged by the
    # check; B = (i, j+1) with gamma 3 then has exactly 4 votes, where
    # counting A's vote (as if A were still good) would give 5 and flag it.
    # C = (i-1, j), D = (i-1, j+1), E = (i+1, j+1) have gamma 3.
    ic, jc = q(.5, n1), q(.5, n2)
    for k in ks:
        flags[ic - 3:ic + 4, jc - 3:jc + 5, k] = 0
        for (i, j, G) in ((ic, jc, 8.0), (ic, jc + 1, 3.0), (ic - 1, jc, 3.0), (ic - 1, jc + 1, 3.0),
                          (ic + 1, jc + 1, 3.0)):
            ii, jj, kk = i + nb, j + nb, k + b3
            p[ii, jj, kk, 2] = np.sqrt(G * G - 1.0) / np.sqrt(gcov11[ii, jj])
            p[ii, jj, kk, 3] = 0.0
            p[ii, jj, kk, 4] = 0.0
    diff = np.flatnonzero(p.view(np.int64) != base.view(np.int64))
    return diff, p.reshape(-1)[diff], flags


def run_fixup(harness, p, flags, weno, finalstep):
    with tempfile.TemporaryDirectory() as tmp:
        fin = os.path.join(tmp, "in.bin")
        with open(fin, "wb") as f:
            f.write(np.ascontiguousarray(p, "<f8").tobytes())
            f.write(np.ascontiguousarray(flags, "<i4").tobytes())
        out = os.path.join(tmp, "out")
        subprocess.run([harness, "fixup", fin, out, str(weno), str(finalstep)], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        return np.load(os.path.join(out, "p_out.npy")), np.load(os.path.join(out, "pflag_out.npy"))


def make_fixup(tag, base_name, seed):
    """post_fixup on a planted state (see plant_fixup), both substeps: the
    changes against the input state are stored sparsely."""
    base = np.load(os.path.join(HERE, base_name))["p_init"]
    rng = np.random.default_rng(seed)
    idx, val, flags = plant_fixup(base, rng)
    p_in = base.copy().reshape(-1)
    p_in[idx] = val
    p_in = p_in.reshape(base.shape)
    out = dict(base=np.array(base_name), mod_idx=idx, mod_val=val, flags=flags, seed=np.int64(seed))
    for fs in (0, 1):
        po, fo = run_fixup(os.path.join(REFDIR, "harness_" + tag), p_in, flags, 0, fs)
        ch = np.flatnonzero(po.view(np.int64) != p_in.view(np.int64))
        out["fs%d_idx" % fs] = ch
        out["fs%d_val" % fs] = po.reshape(-1)[ch]
        out["fs%d_pflag" % fs] = fo.astype(np.int8)
        print(tag, "finalstep", fs, "changed", ch.size, "flags", np.unique(fo, return_counts=True))
    return out


def main_fixup():
    for tag, (n1, n2, n3) in {"t2d": (32, 32, 1), "t3d": (16, 16, 8)}.items():
        subprocess.run([BUILD, tag, str(n1), str(n2), str(n3)], check=True, stdout=subprocess.DEVNULL)
    np.savez_compressed(os.path.join(HERE, "fixup_t2d.npz"), **make_fixup("t2d", "t2d_mc.npz", 31))
    np.savez_compressed(os.path.join(HERE, "fixup_t3d.npz"), **make_fixup("t3d", "t3d_weno.npz", 32))


def run_diag(harness, weno, nsteps):
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "out")
        subprocess.run([harness, "diag", out, str(weno), str(nsteps)], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        d = {f[:-4]: np.load(os.path.join(out, f)) for f in os.listdir(out) if f.endswith(".npy")}
    d["digest_final"] = np.array(state_digest(d.pop("p_final")))
    d["weno"] = np.int32(weno)
    d["nsteps"] = np.int32(nsteps)
    return d


def main_diag():
    """constotal / divbmaxavg / diag_flux of the reference after a few steps
    (the torus fixtures' grids): diag_t2d (MC, 5 steps), diag_t3d (WENO5BND,
    2 steps), diag_t2d_para (PARA + RK4, 3 steps: pcum over every
    substep), diag_t3d_fv (the finite-volume scheme, 2 steps: constotal of
    the carried state)."""
    for tag, (n1, n2, n3) in {"t2d": (32, 32, 1), "t3d": (16, 16, 8)}.items():
        subprocess.run([BUILD, tag, str(n1), str(n2), str(n3)], check=True, stdout=subprocess.DEVNULL)
    np.savez_compressed(os.path.join(HERE, "diag_t2d.npz"), **run_diag(os.path.join(REFDIR, "harness_t2d"), 0, 5))
    np.savez_compressed(os.path.join(HERE, "diag_t3d.npz"), **run_diag(os.path.join(REFDIR, "harness_t3d"), 1, 2))
    np.savez_compressed(os.path.join(HERE, "diag_t2d_para.npz"),
                        **run_diag(os.path.join(REFDIR, "harness_t2d"), 2, 3))
    np.savez_compressed(os.path.join(HERE, "diag_t3d_fv.npz"), **run_diag(os.path.join(REFDIR, "harness_t3d"), 3, 2))


def rdump_fixture(h, scheme):
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.run([h, "rdump", "w", str(scheme), "2"], cwd=tmp, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        text = open(os.path.join(tmp, "dumps", "rdump-0"), "rb").read()
        subprocess.run([h, "rdread", "r", str(scheme)], cwd=tmp, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        pr = np.load(os.path.join(tmp, "r", "p.npy"))
        un = np.load(os.path.join(tmp, "r", "unew.npy"))
        hd = np.load(os.path.join(tmp, "r", "header.npy"))
    nb = (pr.shape[0] - 32) // 2
    return dict(text=np.frombuffer(text, np.uint8), p_active=pr[nb:nb + 32, nb:nb + 32], unew=un, header=hd)


def main_rdump():
    """Restart dumps written by the reference (restart_write after 2 steps of
    the 32x32 torus) and what the reference's own restart_read makes of
    them: the file's bytes, the parsed p and unew of the active zones and the
    header values (t, tf, nstep, a, gam, cour, dt, lim, fluxmethod, R0, Rin,
    Rout, hslope, RHOMIN, UUMIN, GAMMAMAX).  rdump_t2d: MC + LaxF + RK2;
    rdump_t2d_fv: the torus file's finite-volume PARA + HLL + RK4 scheme;
    rdump_t2d_rk4, _rk3, _euler: PARA + HLL + RK4 without the finite-volume
    scheme, MC + LaxF + TVD RK3 and MC + LaxF + Euler."""
    subprocess.run([BUILD, "t2d", "32", "32", "1"], check=True, stdout=subprocess.DEVNULL)
    h = os.path.join(REFDIR, "harness_t2d")
    np.savez_compressed(os.path.join(HERE, "rdump_t2d.npz"), **rdump_fixture(h, 0))
    np.savez_compressed(os.path.join(HERE, "rdump_t2d_fv.npz"), **rdump_fixture(h, 3))
    for name, scheme in (("rk4", 2), ("rk3", 4), ("euler", 5)):
        np.savez_compressed(os.path.join(HERE, "rdump_t2d_%s.npz" % name), **rdump_fixture(h, scheme))


def main_para():
    """The PARA + HLL + RK4 tori (schemes 2 and 3 of the harness)."""
    builds = {"t2d": (32, 32, 1), "t3d": (16, 16, 8)}
    for tag, (n1, n2, n3) in builds.items():
        subprocess.run([BUILD, tag, str(n1), str(n2), str(n3)], check=True, stdout=subprocess.DEVNULL)
    np.savez_compressed(os.path.join(HERE, "t2d_para.npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), 2, 5))
    np.savez_compressed(os.path.join(HERE, "t3d_para.npz"), **run_torus(os.path.join(REFDIR, "harness_t3d"), 2, 1))
    np.savez_compressed(os.path.join(HERE, "t2d_fv.npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), 3, 5))
    np.savez_compressed(os.path.join(HERE, "t3d_fv.npz"), **run_torus(os.path.join(REFDIR, "harness_t3d"), 3, 2))
    np.savez_compressed(os.path.join(HERE, "t2d_rk3.npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), 4, 5))
    np.savez_compressed(os.path.join(HERE, "t2d_euler.npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), 5, 5))
    for scheme, name in ((6, "t2d_force"), (7, "t2d_musta"), (8, "t2d_hlllaxf1")):
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), scheme, 5))
    np.savez_compressed(os.path.join(HERE, "t2d_fvweno.npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), 9, 3))
    np.savez_compressed(os.path.join(HERE, "t3d_fvweno.npz"), **run_torus(os.path.join(REFDIR, "harness_t3d"), 9, 1))
    np.savez_compressed(os.path.join(HERE, "t2d_fv_30.npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), 3, 30))
    np.savez_compressed(os.path.join(HERE, "t3d_weno_10.npz"), **run_torus(os.path.join(REFDIR, "harness_t3d"), 1, 10))
    np.savez_compressed(os.path.join(HERE, "t3d_fv_10.npz"), **run_torus(os.path.join(REFDIR, "harness_t3d"), 3, 10))
    for scheme, name in ((10, "t2d_vanl"), (11, "t2d_minm"), (12, "t2d_donor")):
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run_torus(os.path.join(REFDIR, "harness_t2d"), scheme, 5))


def main():
    if "--para" in sys.argv:
        return main_para()
    if "--rdump" in sys.argv:
        return main_rdump()
    if "--diag" in sys.argv:
        return main_diag()
    if "--long50" in sys.argv:
        return main_long50()
    if "--c5" in sys.argv:
        return main_c5()
    if "--long" in sys.argv:
        return main_long("--full" in sys.argv)
    if "--fixup" in sys.argv:
        return main_fixup()
    builds = {"t2d": (32, 32, 1), "t3d": (16, 16, 8)}
    for tag, (n1, n2, n3) in builds.items():
        subprocess.run([BUILD, tag, str(n1), str(n2), str(n3)], check=True, stdout=subprocess.DEVNULL)
    h2 = os.path.join(REFDIR, "harness_t2d")
    h3 = os.path.join(REFDIR, "harness_t3d")
    np.savez_compressed(os.path.join(HERE, "c2p_g43.npz"), **run_c2p(h2, 4.0 / 3.0, 1024, C2P_SEED))
    np.savez_compressed(os.path.join(HERE, "c2p_g53.npz"), **run_c2p(h2, 5.0 / 3.0, 1024, C2P_SEED + 1))
    np.savez_compressed(os.path.join(HERE, "t2d_mc.npz"), **run_torus(h2, 0, 5))
    np.savez_compressed(os.path.join(HERE, "t2d_weno.npz"), **run_torus(h2, 1, 5))
    np.savez_compressed(os.path.join(HERE, "t3d_weno.npz"), **run_torus(h3, 1, 1))
    main_para()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    sys.exit(main())
