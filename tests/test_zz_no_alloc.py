"""hg_utoprim_batch_noalloc (the batched inversion with its scratch in the
library's static device array) against hg_utoprim_batch on the same inputs:
bit-identical primitives, failure codes, iteration counts and fallback
counts, on two calls in a row, the second with a smaller batch (the array
keeps the first call's counters and list, which the second must not see)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(fn, b, gam, mode):
    from harm_finalwham_amd import harmgpu
    pr = b["p_guess"].clone()
    n = pr.shape[1]
    pflag = torch.full((n,), -7, dtype=torch.int32, device=pr.device)
    iters = torch.full((n,), -7, dtype=torch.int32, device=pr.device)
    fn(b["U"], b["gcov"], b["gcon"], b["gdet"], b["gcovpert"], pr, pflag, iters, gam=gam, mode=mode)
    torch.cuda.synchronize()
    return pr, pflag, iters, harmgpu.c2p_fallbacks()


@pytest.mark.parametrize("mode_name", ["HG_C2P_MIXED", "HG_C2P_FP64"])
def test_noalloc_matches_allocating_entry_point(mode_name):
    from harm_finalwham_amd import harmgpu
    mode = getattr(harmgpu, mode_name)
    gam = 4.0 / 3.0
    for call, (n, seed) in enumerate([(1 << 18, 31), (40001, 32)]):
        print("case", mode_name, "call", call, "n", n, "seed", seed, flush=True)
        b = harmgpu.make_c2p_batch(n, gam=gam, seed=seed)
        pa, fa, ia, nfa = _run(harmgpu.utoprim_batch, b, gam, mode)
        pz, fz, iz, nfz = _run(harmgpu.utoprim_batch_noalloc, b, gam, mode)
        assert torch.equal(fa, fz)
        assert torch.equal(ia, iz)
        assert torch.equal(pa, pz)
        assert int((fz == -7).sum()) == 0  # every zone written
        if mode == harmgpu.HG_C2P_MIXED:
            assert nfa == nfz
            print("fallbacks", nfz, flush=True)


def test_noalloc_rejects_batches_beyond_its_array():
    from harm_finalwham_amd import harmgpu
    import ctypes as C
    t = torch.zeros(16, 1, dtype=torch.float64, device="cuda")
    f = torch.zeros(1, dtype=torch.int32, device="cuda")
    p = lambda x: C.c_void_p(x.data_ptr())
    rc = harmgpu.lib().hg_utoprim_batch_noalloc((1 << 24) + 1, 4.0 / 3.0, p(t), p(t), p(t), p(t), p(t), p(t), p(f),
                                                None, harmgpu.HG_C2P_MIXED, None)
    assert rc != 0
    assert b"exceeds the static scratch" in harmgpu.lib().hg_last_error()
