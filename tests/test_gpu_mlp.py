"""Learned Newton start of the 2D inversion on the GPU.

The matrix-core forward and the vector backward against the CPU FP64
reference of the same network (mlp_ref, itself pinned by finite
differences); training on inversions the library solved; the inversion
from the learned start against the plain FP64 inversion: identical failure
codes, primitives equal to the Newton tolerance, and the prediction closer
to the solution than the plain guess is."""
import numpy as np
import pytest

import mlp_ref

pytestmark = pytest.mark.gpu

GROUPS = {"W1": (0, 512), "b1": (512, 576), "W2": (576, 4672), "b2": (4672, 4736), "W3": (4736, 4800),
          "b3": (4800, 4801)}


def _mlp(w):
    from harm_finalwham_amd import harmgpu
    m = harmgpu.Mlp(seed=1)
    m.set_params(w)
    return m


def test_forward_matches_reference():
    import torch
    rng = np.random.default_rng(1)
    w = mlp_ref.random_params(4).astype(np.float32)
    X = rng.normal(0.0, 1.5, (4099, 8)).astype(np.float32)  # not a multiple of 16 or 64
    m = _mlp(w)
    y = m.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    yr, _ = mlp_ref.forward(w.astype(np.float64), X, quant=True)
    err = np.abs(y - yr)
    scale = max(1.0, np.abs(yr).max())
    print("forward max err", err.max(), "median", np.median(err), "scale", scale, flush=True)
    # FP32 accumulation and the hardware tanh: ~1e-6; a first-layer
    # activation on the other side of a BF16 rounding boundary: ~1e-3
    assert err.max() <= 5e-3 * scale
    assert np.median(err) <= 1e-5 * scale


def test_gradient_matches_reference():
    import torch
    rng = np.random.default_rng(2)
    w = mlp_ref.random_params(6).astype(np.float32)
    X = rng.normal(0.0, 1.5, (3001, 8)).astype(np.float32)
    Y = rng.normal(0.0, 0.5, 3001).astype(np.float32)
    m = _mlp(w)
    g, loss = m.gradient(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda())
    lr, gr = mlp_ref.loss_grad(w.astype(np.float64), X, Y, quant=True)
    assert abs(loss - lr) <= 1e-4 * lr
    for k, (a, b) in GROUPS.items():
        rel = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        print("gradient", k, "relative error", rel, flush=True)
        assert rel <= 1e-3
    assert np.all(g[mlp_ref.NTRAIN:] == 0.0)


def _solve(b, gam, mlp=None, mode=None):
    import torch
    from harm_finalwham_amd import harmgpu
    pr = b["p_guess"].clone()
    n = pr.shape[1]
    pflag = torch.zeros(n, dtype=torch.int32, device=pr.device)
    iters = torch.zeros_like(pflag)
    if mlp is None:
        harmgpu.utoprim_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag, iters, gam=gam,
                              mode=harmgpu.HG_C2P_FP64 if mode is None else mode)
        return pr, pflag, iters, None
    cnt = harmgpu.utoprim_batch_learned(mlp, b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag,
                                        iters, gam=gam)
    return pr, pflag, iters, cnt


def _errors(p, pref, fref, gam):
    """max errors of p against pref over the zones pref solved: rho relative,
    u relative to rho + gam u, velocity relative to max(1, largest component)"""
    ok = fref == 0
    d = (p - pref)[:, ok].abs()
    ref = pref[:, ok].abs()
    wden = ref[0] + gam * ref[1]
    vmag = pref[2:5, ok].abs().amax(dim=0).clamp(min=1.0)
    return [float(x) for x in ((d[0] / ref[0]).max(), (d[1] / wden).max(), (d[2:5].amax(dim=0) / vmag).max())]


def train_on_own_inversions(n, gam, seed, epochs):
    """Train a network on n inversions of the benchmark distribution solved
    by the plain FP64 path; returns the network and the loss history."""
    from harm_finalwham_amd import harmgpu
    import torch
    tr = harmgpu.make_c2p_batch(n, gam=gam, seed=seed)
    p64, f64, _, _ = _solve(tr, gam)
    X, y, ok = harmgpu.mlp_features(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], tr["p_guess"], p_solved=p64,
                                    gam=gam)
    sel = (ok == 1) & (f64 == 0)
    X, y = X[sel].contiguous(), y[sel].contiguous()
    mu, sd = X.mean(0), X.std(0)
    sd[sd < 1e-6] = 1.0
    m = harmgpu.Mlp(seed=7)
    m.set_normalisation(mu.cpu().numpy(), sd.cpu().numpy())
    perm = torch.randperm(X.shape[0], device=X.device, generator=torch.Generator(X.device).manual_seed(seed))
    X, y = X[perm].contiguous(), y[perm].contiguous()
    hist = []
    for ep in range(epochs):
        lr = 2e-3 if ep < epochs // 2 else 5e-4
        hist.append(m.train(X, y, epochs=1, batch=2048, lr=lr))
    return m, hist, float(y.pow(2).mean())


def test_untrained_network_keeps_reference_results():
    """With random weights the residual comparison and the fallback keep
    every failure code, and the solution is the converged root
    (HG_C2P_REFINED) to 1e-10."""
    from harm_finalwham_amd import harmgpu
    gam = 4.0 / 3.0
    b = harmgpu.make_c2p_batch(1 << 14, gam=gam, seed=21)
    m = _mlp(mlp_ref.random_params(3).astype(np.float32))
    p64, f64, i64, _ = _solve(b, gam)
    pre, fre, _, _ = _solve(b, gam, mode=harmgpu.HG_C2P_REFINED)
    pl, fl, il, (nlearn, nfb) = _solve(b, gam, m)
    print("untrained: learned starts", nlearn, "fallbacks", nfb, flush=True)
    import torch
    assert torch.equal(f64, fl)
    assert max(_errors(pl, pre, fre, gam)) <= 1e-10
    assert torch.equal(pl[5:8], p64[5:8])


@pytest.mark.parametrize("gam", [4.0 / 3.0])
def test_learned_start_gain(gam):
    """A network trained on this library's own inversions: its start is
    closer to the root than the plain guess, the mixed-precision iteration
    from it needs fewer steps per wavefront than from the plain start, and
    the result is the converged root to 1e-10 with the reference's codes."""
    import torch
    from harm_finalwham_amd import harmgpu
    m, hist, y2 = train_on_own_inversions(1 << 18, gam, seed=101, epochs=12)
    print("loss by epoch", [round(x, 6) for x in hist], "plain-guess mean square", y2, flush=True)
    assert hist[-1] < hist[0] / 5
    ev = harmgpu.make_c2p_batch(1 << 18, gam=gam, seed=202)  # held out
    p64, f64, i64, _ = _solve(ev, gam)
    pre, fre, _, _ = _solve(ev, gam, mode=harmgpu.HG_C2P_REFINED)
    pmx, fmx, imx, _ = _solve(ev, gam, mode=harmgpu.HG_C2P_MIXED)
    pl, fl, il, (nlearn, nfb) = _solve(ev, gam, m)
    X, y, ok = harmgpu.mlp_features(ev["U"], ev["gcov"], ev["gcon"], ev["gdet"], ev["p_guess"], p_solved=p64, gam=gam)
    s = (ok == 1) & (f64 == 0)
    resid = (m.forward(X) - y)[s]
    rms_plain = float(y[s].pow(2).mean().sqrt())
    rms_learn = float(resid.pow(2).mean().sqrt())
    err = _errors(pl, pre, fre, gam)
    wave_mixed = float(imx.view(-1, 64).amax(dim=1).float().mean())
    wave_learn = float(il.view(-1, 64).amax(dim=1).float().mean())
    print("held out: rms log(W/W_guess) plain %.4g learned %.4g; steps mixed %.3f learned %.3f (fp64 %.3f); "
          "per-wavefront maximum mixed %.3f learned %.3f; learned starts %d fallbacks %d; "
          "max error vs converged root rho, u/w, v %s" % (rms_plain, rms_learn, float(imx.float().mean()),
                                                          float(il.float().mean()), float(i64.float().mean()),
                                                          wave_mixed, wave_learn, nlearn, nfb, err), flush=True)
    # the network's start is far closer to the solution than the plain guess
    # on average (the plain guess is good for most zones and poor for a tail;
    # the residual comparison takes whichever start is better per zone)
    assert rms_learn < rms_plain / 10
    assert nlearn > 0 and nfb < ev["U"].shape[1] // 50
    # a lower per-wavefront maximum of steps (both starts' residuals
    # counted), which sets a wavefront's time
    assert wave_learn < wave_mixed
    assert torch.equal(f64, fl)
    assert max(err) <= 1e-10
    assert torch.equal(pl[5:8], p64[5:8])


def test_learned_entry_rejects_bad_arguments():
    from harm_finalwham_amd import harmgpu
    import ctypes as C
    m = _mlp(mlp_ref.random_params(3).astype(np.float32))
    rc = harmgpu.lib().hg_utoprim_batch_learned(m.h, -1, 4.0 / 3.0, *([None] * 8), None, None)
    assert rc != 0
    assert b"bad arguments" in harmgpu.lib().hg_last_error()
    rc = harmgpu.lib().hg_utoprim_batch_learned(m.h, 0, 4.0 / 3.0, *([None] * 8), None, None)
    assert rc == 0
    w = m.params()
    w[mlp_ref.OFF["sd"]] = 0.0
    with pytest.raises(harmgpu.HarmError):
        m.set_params(w)
    del C


def _torus_dumps(n1, n2, nsteps, seed=7):
    """The final-substep inversions of nsteps RK2 steps of the 2-D torus
    (MC + LaxF), one batch per step (hg_dump_inversions)."""
    from harm_finalwham_amd import harmgpu
    g = harmgpu.Grid(n1, n2, 1, nb=3, lim=harmgpu.HG_LIM_MC, fluxmethod=harmgpu.HG_FLUX_LAXF)
    g.build_mks_geometry()
    dt = g.init_torus(rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=seed)
    out = []
    for _ in range(nsteps):
        dt, _, _ = g.step_ch(dt)
        out.append(g.dump_inversions())
    return g, out


def test_dumped_step_inversions_reproduce_the_step():
    """The batch hg_dump_inversions returns is what the step inverted: the
    FP64 batch inversion of it gives the step's primitives in every zone that
    no floor (fixup.c:1542, rescaletype 0: rho >= 1e-4 r^-1.5, u >= 1e-6
    r^-2.5) or pole row (POLEDEATH) touched afterwards -- bit for bit where
    the zone's metric table is exactly symmetric, to 1e-13 where gcov's
    mirrored entries differ in the last bit (the batch reads the upper
    triangles, the step the full tables)."""
    import torch
    from harm_finalwham_amd import harmgpu
    print("case dump inversions 32x32", flush=True)
    g, dumps = _torus_dumps(32, 32, 3)
    b = dumps[-1]
    n = b["U"].shape[1]
    pr = b["p_guess"].clone()
    pflag = torch.zeros(n, dtype=torch.int32, device="cuda")
    harmgpu.utoprim_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag, None)
    assert (pflag.cpu().numpy() == 0).all()
    p = g.get_prims()[3:-3, 3:-3, 0, :]                      # active zones [i][j][pl]
    pa = pr.cpu().numpy().T.reshape(32, 32, 8).transpose(1, 0, 2)  # t = i + 32 j -> [i][j][pl]
    # the field is copied through both paths
    assert np.array_equal(pa[..., 5:8], p[..., 5:8])
    geo = harmgpu.mks_geometry_tables(32, 32, 1, 3)
    r = geo["r_cent"][3:-3, 3:-3]
    rhoflr = np.maximum(1e-4 * r ** -1.5, 1e-20)
    uuflr = np.maximum(1e-6 * r ** -2.5, 1e-20)
    free = (pa[..., 0] > rhoflr * (1 + 1e-9)) & (pa[..., 1] > uuflr * (1 + 1e-9))
    free[:, 0] = free[:, -1] = False
    gc = geo["gcov"][3:-4, 3:-4, 6]
    sym = np.all(gc == np.swapaxes(gc, -1, -2), axis=(-1, -2))
    same = np.all(pa == p, axis=2)
    scale = np.maximum(np.abs(p[..., :5]), (p[..., 0] + 4.0 / 3.0 * p[..., 1])[..., None])
    rel = (np.abs(pa[..., :5] - p[..., :5]) / scale).max(axis=2)
    print("zones above the floors: %d of %d; symmetric metric among them %d, reproduced bit for bit %d; "
          "largest relative difference elsewhere %.3g" % (free.sum(), free.size, (free & sym).sum(),
                                                         (free & sym & same).sum(), rel[free & ~sym].max()),
          flush=True)
    assert (free & sym).sum() >= 50
    assert same[free & sym].all()
    assert rel[free].max() <= 1e-13


def test_learned_start_on_step_inversions():
    """Trained on the step's own inversions (and the benchmark distribution),
    evaluated on later steps' inversions: the reference's codes, the
    converged root to 1e-10, no more fallbacks than the plain mixed path."""
    import torch
    from harm_finalwham_amd import harmgpu
    from test_gpu_c2p_batch import err3
    print("case learned start on step inversions", flush=True)
    gam = 4.0 / 3.0
    _, dumps = _torus_dumps(64, 64, 8)
    train = dumps[:5]
    Xs, ys = [], []
    for b in train:
        pr = b["p_guess"].clone()
        n = pr.shape[1]
        pflag = torch.zeros(n, dtype=torch.int32, device="cuda")
        harmgpu.utoprim_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag, None, gam=gam)
        X, y, ok = harmgpu.mlp_features(b["U"], b["gcov"], b["gcon"], b["gdet"], b["p_guess"], p_solved=pr, gam=gam)
        sel = (ok == 1) & (pflag == 0)
        Xs.append(X[sel])
        ys.append(y[sel])
    X = torch.cat(Xs).contiguous()
    y = torch.cat(ys).contiguous()
    m = harmgpu.Mlp(seed=7)
    mu, sd = X.mean(0), X.std(0)
    sd[sd < 1e-6] = 1.0
    m.set_normalisation(mu.cpu().numpy(), sd.cpu().numpy())
    hist = [m.train(X, y, epochs=1, batch=2048, lr=2e-3) for _ in range(6)]
    assert hist[-1] <= hist[0]
    for b in dumps[5:]:
        n = b["U"].shape[1]
        res = {}
        for key in ("fp64", "refined", "mixed", "learned"):
            pr = b["p_guess"].clone()
            pflag = torch.zeros(n, dtype=torch.int32, device="cuda")
            it = torch.zeros_like(pflag)
            if key == "learned":
                cnt = harmgpu.utoprim_batch_learned(m, b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr,
                                                    pflag, it, gam=gam)
            else:
                mode = {"fp64": harmgpu.HG_C2P_FP64, "refined": harmgpu.HG_C2P_REFINED,
                        "mixed": harmgpu.HG_C2P_MIXED}[key]
                harmgpu.utoprim_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag, it,
                                      gam=gam, mode=mode)
            res[key] = (pr, pflag, it)
        f64 = res["fp64"][1]
        ok = f64 == 0
        for key in ("mixed", "learned"):
            assert torch.equal(res[key][1], f64)
            e = max(float(x.max()) for x in err3(res[key][0][:, ok], res["refined"][0][:, ok], gam))
            assert e <= 1e-10, (key, e)
        print("steps per zone fp64 %.3f mixed %.3f learned %.3f; learned starts %d fallbacks %d" %
              (float(res["fp64"][2].float().mean()), float(res["mixed"][2].float().mean()),
               float(res["learned"][2].float().mean()), cnt[0], cnt[1]), flush=True)
