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


def _solve(b, gam, mlp=None):
    import torch
    from harm_finalwham_amd import harmgpu
    pr = b["p_guess"].clone()
    n = pr.shape[1]
    pflag = torch.zeros(n, dtype=torch.int32, device=pr.device)
    iters = torch.zeros_like(pflag)
    if mlp is None:
        harmgpu.utoprim_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag, iters, gam=gam)
        return pr, pflag, iters, None
    cnt = harmgpu.utoprim_batch_learned(mlp, b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag,
                                        iters, gam=gam)
    return pr, pflag, iters, cnt


def _errors(p, p64, f64, gam):
    ok = f64 == 0
    d = (p - p64)[:, ok].abs()
    ref = p64[:, ok].abs()
    wden = ref[0] + gam * ref[1]
    vmag = p64[2:5, ok].abs().amax(dim=0)
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
    every failure code and the solution to the Newton tolerance."""
    from harm_finalwham_amd import harmgpu
    gam = 4.0 / 3.0
    b = harmgpu.make_c2p_batch(1 << 14, gam=gam, seed=21)
    m = _mlp(mlp_ref.random_params(3).astype(np.float32))
    p64, f64, i64, _ = _solve(b, gam)
    pl, fl, il, (nlearn, nfb) = _solve(b, gam, m)
    print("untrained: learned starts", nlearn, "fallbacks", nfb, flush=True)
    import torch
    assert torch.equal(f64, fl)
    assert max(_errors(pl, p64, f64, gam)) <= 1e-9
    assert torch.equal(pl[5:8], p64[5:8])


@pytest.mark.parametrize("gam", [4.0 / 3.0])
def test_learned_start_gain(gam):
    import torch
    from harm_finalwham_amd import harmgpu
    m, hist, y2 = train_on_own_inversions(1 << 18, gam, seed=101, epochs=12)
    print("loss by epoch", [round(x, 6) for x in hist], "plain-guess mean square", y2, flush=True)
    assert hist[-1] < hist[0] / 5
    ev = harmgpu.make_c2p_batch(1 << 16, gam=gam, seed=202)  # held out
    p64, f64, i64, _ = _solve(ev, gam)
    pl, fl, il, (nlearn, nfb) = _solve(ev, gam, m)
    X, y, ok = harmgpu.mlp_features(ev["U"], ev["gcov"], ev["gcon"], ev["gdet"], ev["p_guess"], p_solved=p64, gam=gam)
    s = (ok == 1) & (f64 == 0)
    resid = (m.forward(X) - y)[s]
    rms_plain = float(y[s].pow(2).mean().sqrt())
    rms_learn = float(resid.pow(2).mean().sqrt())
    err = _errors(pl, p64, f64, gam)
    wave_plain = float(i64.view(-1, 64).amax(dim=1).float().mean())
    wave_learn = float(il.view(-1, 64).amax(dim=1).float().mean())
    print("held out: rms log(W/W_guess) plain %.4g learned %.4g; evaluations plain %.3f learned %.3f; "
          "per-wavefront maximum plain %.3f learned %.3f; learned starts %d fallbacks %d; "
          "max error rho, u/w, v %s" % (rms_plain, rms_learn, float(i64.float().mean()), float(il.float().mean()),
                                        wave_plain, wave_learn, nlearn, nfb, err), flush=True)
    # the network's start is far closer to the solution than the plain guess
    # on average (the plain guess is good for most zones and poor for a tail;
    # the residual comparison takes whichever start is better per zone)
    assert rms_learn < rms_plain / 10
    assert nlearn > 0 and nfb < ev["U"].shape[1] // 100
    # fewer Newton evaluations (both starts' residuals counted), and a lower
    # per-wavefront maximum, which sets a wavefront's time
    assert float(il.float().mean()) < float(i64.float().mean())
    assert wave_learn < wave_plain
    assert torch.equal(f64, fl)
    assert max(err) <= 1e-9
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
