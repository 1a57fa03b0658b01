#!/usr/bin/env python3
"""Throughput of the per-step GRMHD hot path on one MI355X.

Workload: the magnetised Fishbone-Moncrief torus of init.fishmon.c in 3-D,
256x128x64 zones of modified Kerr-Schild coordinates (a=0.9375, r in
[1.3, 40], hslope 0.3, phi periodic), WENO5BND reconstruction + HLL flux +
flux-CT + Noble 2D inversion + fixup, RK2 at Courant 0.8 (step_ch.c:35),
with the initial data built by the library itself (seeded 2% pressure
perturbation).  All state is resident in HBM before the timed region.

Prints one JSON line: zone-updates/s over K timed steps, the roofline
figure of the dominant kernel (HIP events around every launch of K further
profiled steps, on the stream each kernel runs on; a profile runs the three
directions one after another) and the single-core CPU rate of the
reference's own C build on a bounded sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# algorithmic bytes per zone of each kernel class (unique inputs and outputs
# touched once; geometry is phi independent and amortised over N3)
#   reconstruction + Riemann flux of one direction: read p (8 doubles), write F (8)
#   flux-CT EMFs: read 6 field-flux components, write 3 EMFs
#   update + source + inversion + fixup: read p_i and p_b (16), F1..F3 (24),
#   EMFs (3), write p_f (8)
ALG_BYTES = {"recon+flux": 128, "emf": 72, "update": 408}


def profile_summary(prof, zones):
    """Per logical kernel: total ms, launches, mean ms, algorithmic GB/s."""
    out = {}
    for d in (1, 2, 3):
        r, f = prof.get(f"recon{d}", (0.0, 0)), prof.get(f"flux{d}", (0.0, 0))
        if f[1] == 0:
            continue
        ms = r[0] + f[0]
        out[f"recon+flux{d}"] = dict(total_ms=ms, launches=f[1], bytes_per_zone=ALG_BYTES["recon+flux"])
    for k in ("emf", "update"):
        if prof[k][1]:
            out[k] = dict(total_ms=prof[k][0], launches=prof[k][1], bytes_per_zone=ALG_BYTES[k])
    for v in out.values():
        v["mean_ms"] = v["total_ms"] / v["launches"]
        v["achieved_GBs"] = v["bytes_per_zone"] * zones / (v["mean_ms"] * 1e-3) / 1e9
    out["bound"] = dict(total_ms=prof["bound"][0], launches=prof["bound"][1])
    for k in ("fixup", "prep"):
        if prof.get(k, (0.0, 0))[1]:
            out[k] = dict(total_ms=prof[k][0], launches=prof[k][1])
    return out


# device kernels of each logical kernel (WENO5BND build)
DEVICE_KERNELS = {"recon+flux1": ("hg::k_weno_dpi<1>", "hg::k_weno_x", "hg::k_flux_lr<1>"),
                  "recon+flux2": ("hg::k_weno_dpi<2>", "hg::k_weno_m<2>", "hg::k_flux_lr<2>"),
                  "recon+flux3": ("hg::k_weno_dpi<3>", "hg::k_weno_m<3>", "hg::k_flux_lr<3>"),
                  "emf": ("hg::k_emf",), "update": ("hg::k_update<false, false>", "hg::k_invert<false, 0>")}
TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "hbm_traffic_torus3d_256x128x64_weno_hll.md")
VALU_PROFILE = os.path.join(HERE, "profiles", "valu_issue_torus3d_256x128x64_weno_hll.md")
FP64_PEAK_TFLOPS = 78.6


def fp64_gflop_per_launch(logical):
    """FP64 GFLOP per launch of a logical kernel at 256x128x64 from the
    committed rocprofv3 FP64-instruction summary; None when missing."""
    if not os.path.exists(VALU_PROFILE) or logical not in DEVICE_KERNELS:
        return None
    rows = {}
    for line in open(VALU_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 9 and cells[0].startswith("`"):
            try:
                rows[cells[0].strip("`")] = float(cells[7])
            except ValueError:
                pass
    try:
        return sum(rows[k] for k in DEVICE_KERNELS[logical])
    except KeyError:
        return None


def traffic_per_launch(logical):
    """HBM bytes per launch of a logical kernel at 256x128x64 from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a
    plain copy); None when the summary or a row is missing."""
    if not os.path.exists(TRAFFIC_PROFILE) or logical not in DEVICE_KERNELS:
        return None
    rows = {}
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].startswith("`"):
            try:
                rows[cells[0].strip("`")] = float(cells[4]) * 1e6
            except ValueError:
                pass
    try:
        return sum(rows[k] for k in DEVICE_KERNELS[logical])
    except KeyError:
        return None


def fp64_rate(logical, mean_ms):
    gf = fp64_gflop_per_launch(logical)
    if gf is None:
        return None
    tf = gf / (mean_ms * 1e-3) / 1e3
    return {"achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS,
            "gflop_per_launch": gf, "source": os.path.relpath(VALU_PROFILE, HERE)}


def host_cores():
    """Host cores this process may use, capped at 16 (a one-GPU box's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def reference_timing(scheme, label, timeout_s=300):
    """The reference C build (oracle/_ref/harness_b3d, gcc -O2, single
    threaded) timing one step of the torus at 256x128x8 zones: once alone
    (per core) and then as one process per host core at the same time (the
    aggregate of the box's cores; the reference has no threading of its own)."""
    exe = os.path.join(HERE, "oracle", "_ref", "harness_b3d")
    if not os.path.exists(exe):
        return None

    def launch(tmp):
        return subprocess.Popen([exe, "time", str(scheme), "1"], cwd=tmp, stdout=subprocess.PIPE,
                                stderr=subprocess.DEVNULL, text=True)

    def collect(procs):
        out = []
        for p in procs:
            so, _ = p.communicate(timeout=timeout_s)
            if p.returncode != 0:
                return None
            out.append(json.loads(so.strip().splitlines()[-1]))
        return out

    ncores = host_cores()
    with tempfile.TemporaryDirectory() as tmp:
        dirs = [os.path.join(tmp, str(k)) for k in range(ncores)]
        for d in dirs:
            os.makedirs(d)
        one = collect([launch(dirs[0])])
        if not one:
            return None
        t0 = time.perf_counter()
        many = collect([launch(d) for d in dirs])
        wall = time.perf_counter() - t0
    per_core = one[0]["zone_updates_per_s"]
    res = {"value": per_core, "unit": "zone-updates/s", "cores": 1, "kind": "reference",
           "sample": f"reference harness (gcc -O2 FP64, single thread), {label} at 256x128x8 zones, "
                     f"{one[0]['steps']} step, {one[0]['seconds']:.1f} s timed"}
    if many:
        zu = sum(r["zones"] * r["steps"] for r in many)
        res["aggregate"] = {"value": zu / max(r["seconds"] for r in many), "unit": "zone-updates/s",
                            "cores": ncores, "processes": ncores, "wall_s": wall,
                            "per_process_s": [round(r["seconds"], 2) for r in many]}
    return res


def cpu_baseline():
    """The reference build on the benchmark's scheme (WENO5BND + HLL + RK2)."""
    return reference_timing(1, "WENO5BND + HLL + RK2, the same torus")


def torus_scheme_bench(n1, n2, n3, steps, warmup=2):
    """The torus file's own scheme (init.fishmon.c:96-118): PARA + HLL + RK4
    with the finite-volume step (avgscheme DONOR), on the benchmark grid."""
    import torch
    from harm_finalwham_amd import harmgpu
    g = harmgpu.Grid(n1, n2, n3, nb=3, lim=harmgpu.HG_LIM_PARA, fluxmethod=harmgpu.HG_FLUX_HLL, gam=4.0 / 3.0,
                     cour=0.8, timeorder=4, doenoflux=harmgpu.HG_ENOFINITEVOLUME)
    g.build_mks_geometry(a=0.9375, Rin=1.3, Rout=40.0, R0=0.0, hslope=0.3)
    g.init_torus(rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7)
    shapes = g.autotune(2).as_dict()
    g.step_async(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step_async(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g.profile(True)
    g.step_async(steps)
    torch.cuda.synchronize()
    prof = profile_summary(g.profile_read(), n1 * n2 * n3)
    g.profile(False)
    g.close()
    ms = (t1 - t0) * 1e3 / steps
    return {"workload": f"3-D magnetised torus {n1}x{n2}x{n3}, PARA + HLL + flux-CT + Noble 2D, RK4, finite volume",
            "steps": steps, "ms_per_step": ms, "zone_updates_per_s": n1 * n2 * n3 / (ms * 1e-3),
            "launch_shapes": shapes,
            "kernels_ms_per_step": {k: v["total_ms"] / steps for k, v in prof.items()}}


def torus_scheme_cpu_baseline():
    """The reference build on the torus file's own scheme."""
    return reference_timing(3, "PARA + HLL + RK4 finite volume")


def _c2p_solve(b, gam, mode, reps, mlp=None):
    """Best-of-reps kernel time (HIP events on the launch stream) of one
    batched inversion; returns (pr, pflag, iters, ms, counts)."""
    import torch
    from harm_finalwham_amd import harmgpu
    best, cnt = None, None
    n = b["U"].shape[1]
    pr = b["p_guess"].clone()
    pflag = torch.zeros(n, dtype=torch.int32, device="cuda")
    iters = torch.zeros_like(pflag)
    for _ in range(reps):
        pr.copy_(b["p_guess"])
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        if mlp is None:
            harmgpu.utoprim_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag, iters, gam=gam,
                                  mode=mode)
        else:
            cnt = harmgpu.utoprim_batch_learned(mlp, b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr,
                                                pflag, iters, gam=gam)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    return pr, pflag, iters, best, cnt


def c2p_errors(p, pref, fref, gam):
    """Max errors over the zones pref solved: rho relative, u relative to
    rho + gam u, velocity relative to max(1, the largest component)."""
    ok = fref == 0
    d = (p - pref)[:, ok].abs()
    ref = pref[:, ok].abs()
    return [float((d[0] / ref[0]).max()), float((d[1] / (ref[0] + gam * ref[1])).max()),
            float((d[2:5].amax(dim=0) / ref[2:5].amax(dim=0).clamp(min=1.0)).max())]


# algorithmic bytes per inversion: U 8, the upper triangles of gcov and gcon
# 10 + 10, gdet, the guess's u and u~^i 4 doubles in; the solution 8 doubles,
# the code and the step count out
C2P_BYTES = (8 + 10 + 10 + 1 + 4 + 8) * 8 + 4 + 4


def c2p_bench(n, reps=3):
    """Batched inversions (the second half of the headline metric): the
    benchmark distribution for gam = 4/3 and 5/3 generated on the GPU,
    inverted by the plain FP64 kernel (the reference's arithmetic), by the
    converged-root kernel (HG_C2P_REFINED: the reference's codes, root
    converged with double-double residuals) and by the mixed-precision
    persistent kernel.  Errors are against the converged root."""
    from harm_finalwham_amd import harmgpu
    import torch
    res = {"n": n, "gam": [], "fp64_ms": [], "refined_ms": [], "mixed_ms": [], "flags_identical": True,
           "fallbacks": 0, "max_err_vs_converged": {"fp64": [0.0, 0.0, 0.0], "mixed": [0.0, 0.0, 0.0]},
           "fp64_iters_per_zone": [], "mixed_iters_per_zone": [], "mixed_wavefront_max_iters": []}
    for gam in (4.0 / 3.0, 5.0 / 3.0):
        b = harmgpu.make_c2p_batch(n, gam=gam, seed=2024)
        p64, f64, i64, t64, _ = _c2p_solve(b, gam, harmgpu.HG_C2P_FP64, reps)
        pre, fre, _, tre, _ = _c2p_solve(b, gam, harmgpu.HG_C2P_REFINED, reps)
        pmx, fmx, imx, tmx, _ = _c2p_solve(b, gam, harmgpu.HG_C2P_MIXED, reps)
        res["fallbacks"] += harmgpu.c2p_fallbacks()
        res["gam"].append(gam)
        res["fp64_ms"].append(t64)
        res["refined_ms"].append(tre)
        res["mixed_ms"].append(tmx)
        res["fp64_iters_per_zone"].append(float(i64.double().mean()))
        res["mixed_iters_per_zone"].append(float(imx.double().mean()))
        res["mixed_wavefront_max_iters"].append(float(imx.view(-1, 64).amax(1).double().mean()))
        res["flags_identical"] &= bool(torch.equal(f64, fmx)) and bool(torch.equal(f64, fre))
        for key, p in (("fp64", p64), ("mixed", pmx)):
            e = c2p_errors(p, pre, fre, gam)
            res["max_err_vs_converged"][key] = [max(a, c) for a, c in zip(res["max_err_vs_converged"][key], e)]
        del b, p64, pre, pmx
    tot = 2 * n
    res["fp64_solves_per_s"] = tot / (sum(res["fp64_ms"]) * 1e-3)
    res["refined_solves_per_s"] = tot / (sum(res["refined_ms"]) * 1e-3)
    res["mixed_solves_per_s"] = tot / (sum(res["mixed_ms"]) * 1e-3)
    res["speedup"] = res["mixed_solves_per_s"] / res["fp64_solves_per_s"]
    res["bytes_per_inversion"] = C2P_BYTES
    for key in ("fp64", "mixed", "refined"):
        res[key + "_hbm_frac"] = res[key + "_solves_per_s"] * C2P_BYTES / (HBM_PEAK_GBS * 1e9)
    res["err_note"] = ("max errors over solved zones against the converged root of the same FP64 system "
                       "(HG_C2P_REFINED): rho relative, u relative to rho + gam u, velocity relative to "
                       "max(1, largest component)")
    return res


MLP_FLOP = 2 * (8 * 64 + 64 * 64 + 64 * 1)  # multiply-adds of one forward pass, x2
BF16_DENSE_TFLOPS = 2500.0                  # MI355X dense BF16 matrix peak


def _timed(fn, reps):
    import torch
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        r = fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    return r, best


def step_inversions(n1=128, n2=128, nsteps=100, seed=7):
    """The final-substep inversions of nsteps RK2 steps of the 2-D torus of
    the plain-method configuration (128x128, MC + LaxF + flux-CT), one batch
    per step (hg_dump_inversions): the step's own conserved states, metric
    and guesses (the midpoint primitives, advance.c:193)."""
    from harm_finalwham_amd import harmgpu
    g = harmgpu.Grid(n1, n2, 1, nb=3, lim=harmgpu.HG_LIM_MC, fluxmethod=harmgpu.HG_FLUX_LAXF, gam=4.0 / 3.0,
                     cour=0.8)
    g.build_mks_geometry(a=0.9375, Rin=1.3, Rout=40.0, R0=0.0, hslope=0.3)
    dt = g.init_torus(rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=seed)
    out = []
    for _ in range(nsteps):
        dt, _, _ = g.step_ch(dt)
        out.append(g.dump_inversions())
    g.close()
    return out


def _cat_batches(bs):
    import torch
    return {k: torch.cat([b[k] for b in bs], dim=-1).contiguous() for k in bs[0]}


def _features_solved(b, gam):
    """Features and targets of a batch the plain FP64 path solves."""
    from harm_finalwham_amd import harmgpu
    p64, f64, _, _, _ = _c2p_solve(b, gam, harmgpu.HG_C2P_FP64, 1)
    X, y, ok = harmgpu.mlp_features(b["U"], b["gcov"], b["gcon"], b["gdet"], b["p_guess"], p_solved=p64, gam=gam)
    sel = (ok == 1) & (f64 == 0)
    return X[sel].contiguous(), y[sel].contiguous()


def _eval_modes(b, gam, m, reps):
    """Plain FP64, converged (refined), mixed from the plain start and mixed
    from the learned start on one batch: times, codes, errors against the
    converged root, Newton steps."""
    import torch
    from harm_finalwham_amd import harmgpu
    n = b["U"].shape[1]
    p64, f64, i64, t64, _ = _c2p_solve(b, gam, harmgpu.HG_C2P_FP64, reps)
    pre, fre, _, tre, _ = _c2p_solve(b, gam, harmgpu.HG_C2P_REFINED, reps)
    pmx, fmx, imx, tmx, _ = _c2p_solve(b, gam, harmgpu.HG_C2P_MIXED, reps)
    nfb_mixed = harmgpu.c2p_fallbacks()
    pl, fl, il, tl, cnt = _c2p_solve(b, gam, None, reps, mlp=m)
    wmax = lambda it: float(it[: (n // 64) * 64].view(-1, 64).amax(1).double().mean())
    return {
        "n": n, "fp64_ms": t64, "refined_ms": tre, "mixed_ms": tmx, "learned_ms": tl,
        "fp64_solves_per_s": n / (t64 * 1e-3), "mixed_solves_per_s": n / (tmx * 1e-3),
        "learned_solves_per_s": n / (tl * 1e-3),
        "learned_speedup_vs_fp64": t64 / tl, "learned_speedup_vs_mixed": tmx / tl,
        "flags_identical": bool(torch.equal(f64, fl)) and bool(torch.equal(f64, fmx)),
        "learned_starts": cnt[0], "fallbacks": {"learned": cnt[1], "mixed": nfb_mixed},
        "max_err_vs_converged": {"learned": c2p_errors(pl, pre, fre, gam), "mixed": c2p_errors(pmx, pre, fre, gam),
                                 "fp64": c2p_errors(p64, pre, fre, gam)},
        "steps_per_zone": {"fp64": float(i64.double().mean()), "mixed": float(imx.double().mean()),
                           "learned": float(il.double().mean())},
        "wavefront_max_steps": {"fp64": wmax(i64), "mixed": wmax(imx), "learned": wmax(il)},
    }, p64, f64


def mlp_bench(n_train, n_eval, epochs=8, reps=3):
    """Learned Newton start: a network trained on n_train inversions of the
    benchmark distribution (gam 4/3) plus the plain-method torus's own step
    inversions (steps 1-60 of 100), which the plain FP64 path solved; then
    held-out inversions -- n_eval of the distribution and the torus's steps
    61-100 -- solved plain (FP64), converged (refined), mixed precision from
    the plain start and mixed precision from the learned start; plus the
    matrix-core forward pass alone."""
    import torch
    from harm_finalwham_amd import harmgpu
    gam = 4.0 / 3.0
    res = {"n_train": n_train, "n_eval": n_eval, "epochs": epochs, "gam": gam}
    tr = harmgpu.make_c2p_batch(n_train, gam=gam, seed=4096)
    X, y = _features_solved(tr, gam)
    del tr
    dumps = step_inversions()
    steps_train = _cat_batches(dumps[:60])
    steps_eval = _cat_batches(dumps[60:])
    del dumps
    X1, y1 = _features_solved(steps_train, gam)
    del steps_train
    res["train_samples"] = {"distribution": int(X.shape[0]), "torus_steps": int(X1.shape[0])}
    X = torch.cat([X, X1]).contiguous()
    y = torch.cat([y, y1]).contiguous()
    del X1, y1
    m = harmgpu.Mlp(seed=7)
    hist, t_train = train_mlp(m, X, y, epochs)
    res["train_s"] = t_train
    res["loss_by_epoch"] = hist
    del X, y

    ev = harmgpu.make_c2p_batch(n_eval, gam=gam, seed=8192)
    dist, p64, f64 = _eval_modes(ev, gam, m, reps)
    Xe, ye, oke = harmgpu.mlp_features(ev["U"], ev["gcov"], ev["gcon"], ev["gdet"], ev["p_guess"], p_solved=p64,
                                       gam=gam)
    yp, tf = _timed(lambda: m.forward(Xe), reps)
    s = (oke == 1) & (f64 == 0)
    dist["rms_logW_error_plain_guess"] = float(ye[s].pow(2).mean().sqrt())
    dist["rms_logW_error_learned"] = float((yp - ye)[s].pow(2).mean().sqrt())
    res["distribution"] = dist
    del ev, Xe, ye, yp
    st, p64, f64 = _eval_modes(steps_eval, gam, m, reps)
    Xe, ye, oke = harmgpu.mlp_features(steps_eval["U"], steps_eval["gcov"], steps_eval["gcon"], steps_eval["gdet"],
                                       steps_eval["p_guess"], p_solved=p64, gam=gam)
    s = (oke == 1) & (f64 == 0)
    st["rms_logW_error_plain_guess"] = float(ye[s].pow(2).mean().sqrt())
    st["rms_logW_error_learned"] = float((m.forward(Xe) - ye)[s].pow(2).mean().sqrt())
    st["source"] = "final-substep inversions of steps 61-100 of the 128x128 MC + LaxF torus"
    res["torus_steps"] = st
    n = n_eval
    tflops = MLP_FLOP * n / (tf * 1e-3) / 1e12
    res["forward"] = {"ms": tf, "inferences_per_s": n / (tf * 1e-3), "flop_per_inference": MLP_FLOP,
                      "roofline": {"bound": "mfma", "achieved": tflops, "peak": BF16_DENSE_TFLOPS,
                                   "unit": "TFLOP/s", "frac": tflops / BF16_DENSE_TFLOPS}}
    return res, m


def train_mlp(m, X, y, epochs, seed=1):
    """Normalise the features, shuffle, and train with Adam (minibatch 2048,
    learning rate 2e-3 then 5e-4); returns (loss by epoch, seconds)."""
    import torch
    mu, sd = X.mean(0), X.std(0)
    sd[sd < 1e-6] = 1.0
    m.set_normalisation(mu.cpu().numpy(), sd.cpu().numpy())
    perm = torch.randperm(X.shape[0], device="cuda", generator=torch.Generator("cuda").manual_seed(seed))
    X, y = X[perm].contiguous(), y[perm].contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    for ep in range(epochs):
        hist.append(m.train(X, y, epochs=1, batch=2048, lr=2e-3 if ep < epochs // 2 else 5e-4))
    torch.cuda.synchronize()
    return hist, time.perf_counter() - t0


def c2p_cpu_baseline(n=1 << 22):
    """The oracle's Noble 2D inversion (C, gcc -O2, one thread; bit-identical
    to the reference's Utoprimgen) on n inversions of the same distribution,
    and the same split over one thread per host core (the C loop releases
    the interpreter, so the threads run in parallel)."""
    import numpy as np
    import torch
    from harm_finalwham_amd import harmgpu
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle
    gam = 4.0 / 3.0
    b = harmgpu.make_c2p_batch(n, gam=gam, seed=2025)
    h = {k: v.cpu().numpy() for k, v in b.items()}
    gpert = np.zeros((n, 4))
    gpert[:, 0] = h["gcovpert"]
    args = (h["gcov"].T.reshape(n, 4, 4).copy(), h["gcon"].T.reshape(n, 4, 4).copy(), h["gdet"].copy(), gpert,
            h["U"].T.copy(), gam, h["p_guess"].T.copy())
    t0 = time.perf_counter()
    pyoracle.c2p_batch(*args)
    dt = time.perf_counter() - t0
    res = {"value": n / dt, "unit": "inversions/s", "cores": 1, "kind": "port",
           "sample": f"oracle Noble 2D (C, gcc -O2, bit-identical to the reference) on {n} inversions, {dt:.2f} s"}
    ncores = host_cores()
    if ncores > 1:
        from concurrent.futures import ThreadPoolExecutor
        parts = np.array_split(np.arange(n), ncores)
        sub = [tuple(a[idx] if isinstance(a, np.ndarray) else a for a in args) for idx in parts]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(ncores) as ex:
            list(ex.map(lambda a: pyoracle.c2p_batch(*a), sub))
        dta = time.perf_counter() - t0
        res["aggregate"] = {"value": n / dta, "unit": "inversions/s", "cores": ncores, "threads": ncores}
    return res


def batched_bench(ntori=48, nsteps=10, seed=5, mlp=None):
    """48 independent 2-D tori of mixed sizes, spins a ~ U[0, 0.98] and
    pressure-maximum radii r_pmax ~ U[10, 14] (WENO5BND + HLL + flux-CT +
    Noble 2D): nsteps RK2 steps of all of them in one HIP graph launch, with
    launch shapes autotuned once per grid size, against the same grids stepped
    one after the other; every grid is compared bit for bit.  First with the
    reference's FP64 inversion, then the same grids in a grouped graph (one
    launch per kernel over all of them) with the FP64 and with the
    mixed-precision inversion, then, with a trained network mlp, with the
    mixed-precision inversion from the learned start (hg_set_inversion) in a
    new graph of the same grids."""
    import math
    import numpy as np
    import torch
    from harm_finalwham_amd import harmgpu
    rng = np.random.default_rng(seed)
    sizes = [(64, 64), (96, 128), (128, 128), (192, 192), (256, 256)]
    specs = [(sizes[int(rng.integers(len(sizes)))], float(rng.uniform(0.0, 0.98)), float(rng.uniform(10.0, 14.0)),
              int(rng.integers(1, 1 << 30))) for _ in range(ntori)]

    def make(spec):
        (n1, n2), a, rmax, sd = spec
        g = harmgpu.Grid(n1, n2, 1, nb=6, lim=harmgpu.HG_LIM_WENO5BND, fluxmethod=harmgpu.HG_FLUX_HLL)
        rh = 1.0 + math.sqrt(1.0 - a * a)
        g.build_mks_geometry(a=a, Rin=0.9 * rh, Rout=40.0, R0=0.0, hslope=0.3)
        g.init_torus(rin=6.0, rmax=rmax, kappa=1e-3, beta=100.0, randfact=4e-2, seed=sd)
        return g

    t0 = time.perf_counter()
    batch = [make(sp) for sp in specs]
    alone = [make(sp) for sp in specs]
    t_setup = time.perf_counter() - t0
    # autotune once per grid size, share the shapes
    tuned = {}
    for g, sp in zip(batch, specs):
        if sp[0] not in tuned:
            tuned[sp[0]] = g.autotune(nsteps=2)
        g.set_tune(tuned[sp[0]])
    for g, sp in zip(alone, specs):
        g.set_tune(tuned[sp[0]])
    zones = sum(n1 * n2 for (n1, n2), _, _, _ in specs)
    graph = harmgpu.Graph(batch, nsteps)
    graph.launch()  # warm-up replay, repeated on the single grids below
    for g in alone:
        g.step_async(nsteps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graph.launch()
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    for g in alone:
        g.step_async(nsteps)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    identical = all(gb.get_dt() == ga.get_dt() and np.array_equal(gb.get_prims(), ga.get_prims())
                    for gb, ga in zip(batch, alone))
    graph.close()

    def grouped_run(label):
        # one launch per kernel over all the grids (hg_graph_create_grouped),
        # with the segment counts it picks for the batch (the per-size tuned
        # ones are for grids that run alone)
        saved = [g.tune() for g in batch]
        for g, t in zip(batch, saved):
            u = g.tune()
            u.weno_nseg = u.weno_nseg3 = 0
            g.set_tune(u)
        gg = harmgpu.Graph(batch, nsteps, grouped=True)
        for g, t in zip(batch, saved):
            g.set_tune(t)
        gg.launch()
        for g in alone:
            g.step_async(nsteps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gg.launch()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        for g in alone:
            g.step_async(nsteps)
        torch.cuda.synchronize()
        same = all(gb.get_dt() == ga.get_dt() and np.array_equal(gb.get_prims(), ga.get_prims())
                   for gb, ga in zip(batch, alone))
        gg.close()
        return {"inversion": label, "graph_zone_updates_per_s": zones * nsteps / t,
                "speedup_vs_fp64_branch_graph": t_graph / t, "bitwise_identical_to_single_runs": same}

    grouped = grouped_run("the reference's FP64 iteration (HG_C2P_FP64)")
    for g in batch + alone:
        g.set_inversion(harmgpu.HG_C2P_MIXED)
    grouped_mixed = grouped_run("mixed precision (HG_C2P_MIXED)")
    learned = None
    if mlp is not None:
        for g in batch + alone:
            g.set_inversion(harmgpu.HG_C2P_LEARNED, mlp)
        graph = harmgpu.Graph(batch, nsteps)
        graph.launch()
        for g in alone:
            g.step_async(nsteps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        graph.launch()
        torch.cuda.synchronize()
        t_lg = time.perf_counter() - t0
        t0 = time.perf_counter()
        for g in alone:
            g.step_async(nsteps)
        torch.cuda.synchronize()
        t_ls = time.perf_counter() - t0
        same = all(gb.get_dt() == ga.get_dt() and np.array_equal(gb.get_prims(), ga.get_prims())
                   for gb, ga in zip(batch, alone))
        graph.close()
        learned = {"inversion": "mixed precision from the learned start (HG_C2P_LEARNED)",
                   "graph_zone_updates_per_s": zones * nsteps / t_lg,
                   "sequential_zone_updates_per_s": zones * nsteps / t_ls,
                   "graph_speedup": t_ls / t_lg, "vs_fp64_inversion_graph": t_graph / t_lg,
                   "bitwise_identical_to_single_runs": same,
                   "grouped": grouped_run("mixed precision from the learned start (HG_C2P_LEARNED)")}
    res = {"tori": ntori, "steps": nsteps, "zones": zones,
           "sizes": {f"{a}x{b}": sum(1 for sp in specs if sp[0] == (a, b)) for a, b in sizes},
           "graph_zone_updates_per_s": zones * nsteps / t_graph,
           "sequential_zone_updates_per_s": zones * nsteps / t_seq,
           "graph_speedup": t_seq / t_graph, "bitwise_identical_to_single_runs": identical,
           "tuned_shapes": {f"{k[0]}x{k[1]}": v.as_dict() for k, v in tuned.items()},
           "setup_s": t_setup, "grouped": grouped, "grouped_mixed_inversion": grouped_mixed,
           "learned_inversion": learned}
    for g in batch + alone:
        g.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--zones", default="256,128,64", help="n1,n2,n3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--serial", action="store_true",
                    help="run the three directions one after another (for kernel traces)")
    ap.add_argument("--c2p-n", type=int, default=1 << 24, help="inversions per gam for the con2prim leg (0: skip)")
    ap.add_argument("--batch-tori", type=int, default=48, help="tori of the batched-graph leg (0: skip)")
    ap.add_argument("--mlp-train-n", type=int, default=1 << 22, help="training inversions of the learned start")
    ap.add_argument("--mlp-eval-n", type=int, default=1 << 24, help="held-out inversions (0: skip the leg)")
    ap.add_argument("--torus-scheme-steps", type=int, default=10,
                    help="steps of the torus file's own scheme (PARA + RK4 + finite volume) leg (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied {a.numel() * 8} bytes three times", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch
    from harm_finalwham_amd import harmgpu

    n1, n2, n3 = (int(x) for x in args.zones.split(","))
    zones = n1 * n2 * n3
    torch.cuda.init()
    t_setup = time.perf_counter()
    g = harmgpu.Grid(n1, n2, n3, nb=6, lim=harmgpu.HG_LIM_WENO5BND, fluxmethod=harmgpu.HG_FLUX_HLL,
                     gam=4.0 / 3.0, cour=0.8)
    g.build_mks_geometry(a=0.9375, Rin=1.3, Rout=40.0, R0=0.0, hslope=0.3)
    dt0 = g.init_torus(rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7)
    t_setup = time.perf_counter() - t_setup
    print(f"# set-up {t_setup:.1f} s, dt0 = {dt0:.6g}", file=sys.stderr, flush=True)
    # launch shapes for this grid (hg_autotune times a few steps from the
    # current state and restores it; shapes never change a result bit)
    t = g.tune() if args.no_autotune else g.autotune(4)
    if args.serial:
        # the three directions one after another, as the HIP-event profile
        # pass below runs them (a kernel trace then sees each kernel alone)
        t.concurrent = 0
        g.set_tune(t)
    shapes = t.as_dict()
    print(f"# launch shapes {json.dumps(shapes)}", file=sys.stderr, flush=True)

    g.step_async(args.warmup)
    torch.cuda.synchronize()
    lib = harmgpu.lib()
    import ctypes as C
    nf, it = C.c_int64(), C.c_int64()
    harmgpu._check(lib.hg_get_counters(g.h, C.byref(nf), C.byref(it)))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step_async(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    harmgpu._check(lib.hg_get_counters(g.h, C.byref(nf), C.byref(it)))
    dt_end = g.get_dt()
    # the per-kernel figures: K more steps with HIP events around every launch
    # (a profile runs the three directions one after another)
    g.profile(True)
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    g.step_async(args.steps)
    torch.cuda.synchronize()
    tp1 = time.perf_counter()
    prof = g.profile_read()
    g.profile(False)

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = zones / (ms_per_step * 1e-3)
    summ = profile_summary(prof, zones)
    kernels = {k: v for k, v in summ.items() if "bytes_per_zone" in v}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    dk = kernels[dom]
    print(f"# {args.steps} steps: {ms_per_step:.3f} ms/step, failed inversions {nf.value}, "
          f"Newton iterations {it.value}, dt {dt0:.6g} -> {dt_end:.6g}", file=sys.stderr, flush=True)
    for k, v in summ.items():
        print(f"#   {k:12s} {json.dumps(v)}", file=sys.stderr, flush=True)

    c2p = None
    if args.c2p_n > 0:
        g.close()
        c2p = c2p_bench(args.c2p_n)
        if not args.no_cpu_baseline:
            c2p["cpu_baseline"] = c2p_cpu_baseline()
        print(f"# con2prim: {json.dumps(c2p)}", file=sys.stderr, flush=True)
    learned, net = None, None
    if args.mlp_eval_n > 0:
        if args.c2p_n <= 0:
            g.close()
        learned, net = mlp_bench(args.mlp_train_n, args.mlp_eval_n)
        print(f"# learned start: {json.dumps(learned)}", file=sys.stderr, flush=True)
    torus_scheme = None
    if args.torus_scheme_steps > 0:
        g.close()
        torus_scheme = torus_scheme_bench(n1, n2, n3, args.torus_scheme_steps)
        if not args.no_cpu_baseline:
            torus_scheme["cpu_baseline"] = torus_scheme_cpu_baseline()
        print(f"# torus scheme: {json.dumps(torus_scheme)}", file=sys.stderr, flush=True)
    batched = None
    if args.batch_tori > 0:
        batched = batched_bench(args.batch_tori, mlp=net)
        print(f"# batched: {json.dumps(batched)}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    out = {
        "metric": "zone-updates/s full GRMHD step",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Fishbone-Moncrief torus built by the library (init.fishmon.c)",
        "config": {"workload": f"3-D magnetised torus {n1}x{n2}x{n3}, WENO5BND + HLL + flux-CT + Noble 2D, RK2",
                   "zones": zones, "global_batch": 1, "seq_len": 0, "parallelism": "none"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dk["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dk["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": traffic_per_launch(dom) if (n1, n2, n3) == (256, 128, 64) else None,
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, HERE),
                     "algorithmic_bytes_per_launch": dk["bytes_per_zone"] * zones,
                     "bytes_per_zone": dk["bytes_per_zone"], "mean_launch_ms": dk["mean_ms"]},
        # the same kernel against the FP64 issue rate that actually bounds it
        "fp64_valu": fp64_rate(dom, dk["mean_ms"]) if (n1, n2, n3) == (256, 128, 64) else None,
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: v["total_ms"] / args.steps for k, v in summ.items()},
        "profiled_ms_per_step": (tp1 - tp0) * 1e3 / args.steps,
        "launch_shapes": shapes,
        "con2prim": c2p,
        "learned_start": learned,
        "batched_tori": batched,
        "torus_scheme": torus_scheme,
        "failed_inversions": nf.value,
        "newton_iterations_per_zone_update": it.value / (2.0 * zones * args.steps),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
