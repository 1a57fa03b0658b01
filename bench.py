#!/usr/bin/env python3
"""Throughput of the per-step GRMHD hot path on one MI355X.

Workload: the magnetised Fishbone-Moncrief torus of init.fishmon.c in 3-D,
256x128x64 zones of modified Kerr-Schild coordinates (a=0.9375, r in
[1.3, 40], hslope 0.3, phi periodic), WENO5BND reconstruction + HLL flux +
flux-CT + Noble 2D inversion + fixup, RK2 at Courant 0.8 (step_ch.c:35),
with the initial data built by the library itself (seeded 2% pressure
perturbation).  All state is resident in HBM before the timed region.

Prints one JSON line: zone-updates/s over K timed steps, the roofline
figure of the dominant kernel (HIP events around every launch of the timed
steps, on the stream the kernels run on) and the single-core CPU rate of
the reference's own C build on a bounded sample of the same workload.
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
    return out


def cpu_baseline(timeout_s=240):
    """Reference C build (oracle/_ref/harness_b3d, gcc -O2, one core) on the
    same torus with phi reduced to 8 zones: 256x128x8 = 262144 zones, 1 step."""
    exe = os.path.join(HERE, "oracle", "_ref", "harness_b3d")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([exe, "time", "1", "1"], cwd=tmp, capture_output=True, text=True, timeout=timeout_s)
    if r.returncode != 0:
        return None
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": res["zone_updates_per_s"], "unit": "zone-updates/s", "cores": 1, "kind": "reference",
            "sample": "reference harness (gcc -O2 FP64, single thread), same torus and scheme at 256x128x8 "
                      f"zones, {res['steps']} RK2 step, {res['seconds']:.1f} s timed"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--zones", default="256,128,64", help="n1,n2,n3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
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

    g.step_async(args.warmup)
    torch.cuda.synchronize()
    lib = harmgpu.lib()
    import ctypes as C
    nf, it = C.c_int64(), C.c_int64()
    harmgpu._check(lib.hg_get_counters(g.h, C.byref(nf), C.byref(it)))

    g.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step_async(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = g.profile_read()
    g.profile(False)
    harmgpu._check(lib.hg_get_counters(g.h, C.byref(nf), C.byref(it)))
    dt_end = g.get_dt()

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
                     "unit": "GB/s", "frac": dk["achieved_GBs"] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_zone": dk["bytes_per_zone"], "mean_launch_ms": dk["mean_ms"]},
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: v["total_ms"] / args.steps for k, v in summ.items()},
        "failed_inversions": nf.value,
        "newton_iterations_per_zone_update": it.value / (2.0 * zones * args.steps),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
